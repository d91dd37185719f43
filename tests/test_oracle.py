"""Pins the FP64 oracle (oracle/ct_oracle.cpp) without the reference.

The reference ships no tests, golden vectors or fixtures, and its CUDA-only
sources cannot be built here, so the oracle is pinned by what the mathematics
fixes:
  * an independent all-pairs interval-intersection implementation (own
    ray/axis intersections, no merge, no chunking, no start-pixel search);
  * adjoint identities <Ax, y> = <x, A^T y> (projection / projection_t,
    backprojection_t / backprojection for ED and parallel, the exact adjoint for
    EA) and the closed form EA backprojection_t = diag(dImg / c_k) B^T;
  * known answers: parallel beam at theta = 0 and pi/2 with dDet = dImg gives
    column / row sums; parallel-beam mass conservation;
  * invariance to the 256-bin detector chunking of the reference's grid;
  * textbook properties of the RL taps and of the conv2d cross-correlation;
  * FBP of a uniform disk reconstructs the disk (all three geometries), and the
    fbp pre-weights equal the cosine of each bin-centre ray measured from the
    ray geometry itself.
Each check is chosen so that a dropped factor, a wrong sign or index, a
swapped axis or a reversed bin order fails it.
"""
import math

import numpy as np
import pytest

from ctlib_amd import phantom
from oracle import oracle as O


def fan(st, V=24, D=17, W=12, H=10, dImg=0.1, dDet=None, Ang0=0.1, dAng=None, s2r=3.0, d2r=2.0, bs=0.02):
    if dDet is None:
        dDet = 0.03 if st == 1 else 0.13
    if dAng is None:
        dAng = 2 * math.pi / V
    return np.array([V, D, W, H, dImg, dDet, Ang0, dAng, s2r, d2r, bs, st], dtype=np.float64)


def para(V=24, D=17, W=12, H=10, dImg=0.1, dDet=0.13, Ang0=0.1, dAng=None, bs=0.02):
    if dAng is None:
        dAng = 2 * math.pi / V
    return np.array([V, D, W, H, dImg, dDet, Ang0, dAng, bs, 2], dtype=np.float64)


GEOMS = {
    "ed": fan(0),
    "ea": fan(1),
    "para": para(),
    "ed_odd_even": fan(0, V=36, D=16, W=9, H=14, Ang0=-0.4, bs=-0.07),
    "ea_wide": fan(1, V=20, D=31, W=16, H=16, dDet=0.05, s2r=4.0, d2r=1.0, bs=0.01),
    # views exactly on pi/4 multiples (axis ties) and reversed quadrants
    "ed_ties": fan(0, V=8, D=21, W=10, H=10, Ang0=0.0, dAng=math.pi / 4),
    "para_ties": para(V=8, D=19, W=10, H=10, Ang0=0.0, dAng=math.pi / 4, bs=0.0),
}


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def rand_in(op, o, B, seed):
    g = O.parse(o)
    rng = np.random.default_rng(seed)
    if op in ("projection", "backprojection_t", "backprojection_t_true"):
        return rng.random((B, g["height"], g["width"]))
    return rng.random((B, g["views"], g["dets"]))


@pytest.mark.parametrize("name", list(GEOMS))
@pytest.mark.parametrize("op", ["projection", "projection_t", "backprojection", "backprojection_t",
                                "backprojection_t_true"])
def test_merge_matches_brute_force(name, op):
    o = GEOMS[name]
    x = rand_in(op, o, 2, 1)
    a = O.apply(op, o, x)
    b = O.brute(op, o, x)
    assert rel(a, b) < 1e-13, (name, op, rel(a, b))


@pytest.mark.parametrize("name", list(GEOMS))
def test_backprojection_sv_matches_brute_force(name):
    o = GEOMS[name]
    y = rand_in("backprojection", o, 2, 11)
    a = O.apply("backprojection_sv", o, y)
    b = O.brute("backprojection_sv", o, y)
    assert a.shape == (2, y.shape[1]) + b.shape[2:]
    assert rel(a, b) < 1e-13, (name, rel(a, b))


@pytest.mark.parametrize("name", ["ed", "ea", "para", "ed_ties"])
def test_backprojection_sv_is_per_view_backprojection(name):
    """fan_ed_kernel.cu:945-1161 is bprj_fan_ed (:707-923) writing image[channel][view]:
    image v is the backprojection of the sinogram with every other view zeroed,
    and the views sum to the backprojection."""
    o = GEOMS[name]
    g = O.parse(o)
    y = rand_in("backprojection", o, 2, 12)
    sv = O.apply("backprojection_sv", o, y)
    full = O.apply("backprojection", o, y)
    np.testing.assert_allclose(sv.sum(axis=1), full, rtol=1e-12, atol=1e-13 * np.abs(full).max())
    for v in (0, 3, g["views"] - 1):
        ym = np.zeros_like(y)
        ym[:, v] = y[:, v]
        np.testing.assert_allclose(sv[:, v], O.apply("backprojection", o, ym), rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("name", list(GEOMS))
def test_adjoint_identities(name):
    o = GEOMS[name]
    g = O.parse(o)
    rng = np.random.default_rng(2)
    x = rng.random((2, g["height"], g["width"]))
    y = rng.random((2, g["views"], g["dets"]))
    lhs, rhs = (O.apply("projection", o, x) * y).sum(), (x * O.apply("projection_t", o, y)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    lhs, rhs = (O.apply("backprojection_t_true", o, x) * y).sum(), (x * O.apply("backprojection", o, y)).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    if g["scan_type"] != 1:
        # backprojection_t of the reference is the exact adjoint for ED and parallel beam
        np.testing.assert_allclose(O.apply("backprojection_t", o, x), O.apply("backprojection_t_true", o, x),
                                   rtol=0, atol=1e-14 * np.abs(O.apply("backprojection_t", o, x)).max())


def test_ea_backprojection_t_quirk_closed_form():
    """fan_ea_kernel.cu:590-607: EA backprojection_t = diag(dImg / c_k) * (exact adjoint of B)."""
    o = GEOMS["ea"]
    g = O.parse(o)
    x = np.random.default_rng(4).random((1, g["height"], g["width"]))
    bt = O.apply("backprojection_t", o, x)[0]
    bt_true = O.apply("backprojection_t_true", o, x)[0]
    for v in range(g["views"]):
        _, rev, _, c = O.view(o, v)
        fac = g["dImg"] / c
        if rev:
            fac = fac[::-1]
        np.testing.assert_allclose(bt[v], fac * bt_true[v], rtol=1e-13, atol=1e-300)
    assert not np.allclose(bt, bt_true)


@pytest.mark.parametrize("name", ["ed", "ea", "para", "ed_odd_even"])
def test_chunk_invariance(name):
    o = GEOMS[name]
    for op in ("projection", "projection_t", "backprojection", "backprojection_t"):
        x = rand_in(op, o, 1, 5)
        ref = O.apply(op, o, x, chunk=256)
        for chunk in (1, 4, 7):
            np.testing.assert_allclose(O.apply(op, o, x, chunk=chunk), ref, rtol=1e-12,
                                       atol=1e-13 * np.abs(ref).max())


def test_parallel_known_answers():
    """theta = 0: bin j = column j, y = dImg * column sum; theta = pi/2 gives row sums."""
    n = 9
    x = np.random.default_rng(6).random((1, n, n))
    o = para(V=2, D=n, W=n, H=n, dImg=0.25, dDet=0.25, Ang0=0.0, dAng=math.pi / 2, bs=0.0)
    y = O.apply("projection", o, x)[0]
    np.testing.assert_allclose(y[0], 0.25 * x[0].sum(axis=0), rtol=1e-13)
    # theta = pi/2: the detector axis points along +y; bin j collects image row n-1-j
    np.testing.assert_allclose(y[1], 0.25 * x[0].sum(axis=1)[::-1], rtol=1e-13)


def test_parallel_conservation():
    """Every view of a parallel scan carries the whole mass: sum_d y * dDet = dImg^2 * sum x."""
    o = para(V=13, D=40, W=16, H=16, dImg=0.1, dDet=0.07, Ang0=0.2, dAng=math.pi / 13, bs=0.01)
    x = np.random.default_rng(8).random((1, 16, 16))
    y = O.apply("projection", o, x)[0]
    np.testing.assert_allclose(y.sum(axis=1) * 0.07, 0.01 * x.sum(), rtol=1e-12)


def test_fan_mass_relation():
    """Fan-ED with a uniform image: each bin value is the chord length (times the
    image value) of its central ray through the image square, to first order."""
    o = fan(0, V=4, D=64, W=32, H=32, dImg=1 / 16, dDet=0.01, Ang0=0.0, dAng=math.pi / 2, s2r=8.0, d2r=4.0,
            bs=0.0)
    x = np.ones((1, 32, 32))
    y = O.apply("projection", o, x)[0]
    # central bins: ray through the isocentre crosses the 2x2 square with length 2
    D = 64
    mid = (y[:, D // 2 - 1] + y[:, D // 2]) / 2
    np.testing.assert_allclose(mid, 2.0, rtol=2e-3)


def test_rl_taps_frequency_response():
    """Kak & Slaney: the DTFT of the RL taps is the ramp |f| on the band |f| <= 1/(2 delta)."""
    delta = 0.5
    D = 4001
    h = O.rl_taps(D, delta)
    n = np.arange(-(D - 1), D)
    for f in (0.1, 0.3, 0.55, 0.8):
        H = delta * np.sum(h * np.cos(2 * math.pi * f * n * delta))
        assert abs(H - f) < 2e-4, (f, H)
    # the taps are even, non-positive off-centre and vanish at even offsets
    assert np.allclose(h, h[::-1])
    assert np.all(h[D - 1 + 2::2] == 0) and np.all(h[D:] <= 0)


def test_ramp_is_cross_correlation():
    """conv2d(q, h, padding=D-1) is the 'same' cross-correlation of the pre-weighted row."""
    for st, o in (("ed", GEOMS["ed"]), ("ea", GEOMS["ea"]), ("para", GEOMS["para"])):
        g = O.parse(o)
        y = np.random.default_rng(10).random((3, g["dets"]))
        pw, delta = O.fbp_preweight(o)
        h = O.rl_taps(g["dets"], delta)
        q = O.ramp(o, y)
        for r in range(3):
            full = np.correlate(y[r] * pw, h, mode="full")  # c[k] = sum_n a[n+k-(N-1)] * h[n]
            D = g["dets"]
            assert full.shape[0] == 3 * D - 2
            # q[j] = sum_m a[m] h[m - j + D - 1] = c[k] at lag k = j - (D - 1), i.e. full[j + D - 1]
            np.testing.assert_allclose(q[r], full[D - 1:2 * D - 1], rtol=1e-12, atol=1e-12 * np.abs(full).max())
            # the same from the definition of a zero-padded convolution with the flipped taps
            conv = np.convolve(y[r] * pw, h[::-1], mode="full")
            np.testing.assert_allclose(q[r], conv[D - 1:2 * D - 1], rtol=1e-12, atol=1e-12 * np.abs(conv).max())


def test_preweights():
    o = fan(0, D=8, dDet=0.2, s2r=3.0, d2r=1.0)
    pw, delta = O.fbp_preweight(o)
    vd = 0.2 * 3 / 4
    assert math.isclose(delta, vd, rel_tol=1e-15)
    u = (np.arange(8) - 4 + 0.5) * vd
    # cosine of the ray angle on the virtual detector, times the bin width
    np.testing.assert_allclose(pw, vd * np.cos(np.arctan(u / 3.0)), rtol=1e-14)
    o = fan(1, D=8, dDet=0.04, s2r=3.0)
    pw, delta = O.fbp_preweight(o)
    np.testing.assert_allclose(pw, 3.0 * 0.04 * np.cos((np.arange(8) - 3.5) * 0.04), rtol=1e-14)


def _bin_centre_angles(o, v):
    """Signed angle between the central ray (source -> isocentre) and the ray through the centre of
    every bin of view v (output bin order), from the oracle's axis-mapped bin edges and the source
    position: each edge point on the sweep axis gives an edge ray, and a fan-EA bin centre lies half
    way in angle between its two edge rays."""
    g = O.parse(o)
    xmode, rev, e, _ = O.view(o, v)
    th = v * g["dAng"] + g["Ang0"]
    S = np.array([-math.sin(th), math.cos(th)]) * g["s2r"]
    c = -S
    P = np.stack([e, np.zeros_like(e)], 1) if xmode else np.stack([np.zeros_like(e), e], 1)
    r = P - S
    a = np.arctan2(c[0] * r[:, 1] - c[1] * r[:, 0], r @ c)
    mid = 0.5 * (a[1:] + a[:-1])
    return mid[::-1] if rev else mid


def test_ea_preweight_geometry():
    """fan_ea_kernel.cu:1075-1080: the EA fbp pre-weight is s2r * dDet * cos(gamma_j), gamma_j the angle
    of bin j's centre ray off the central ray.  gamma_j is measured here from the ray geometry (source
    position and the oracle's mapped bin edges), not from the (j - D/2 + 1/2) dDet formula, and the
    angles themselves must be D/2 - 1/2 bins apart from the centre on both sides (no half-bin shift)."""
    for D, dDet, s2r in ((17, 0.03, 3.0), (24, 0.011, 5.95), (31, 0.05, 4.0)):
        o = fan(1, V=12, D=D, W=10, H=10, dImg=0.1, dDet=dDet, Ang0=0.3, s2r=s2r, d2r=2.0, bs=0.0)
        pw, delta = O.fbp_preweight(o)
        assert delta == pytest.approx(dDet, rel=1e-15)
        for v in range(12):
            gam = _bin_centre_angles(o, v)
            np.testing.assert_allclose(np.abs(gam), np.abs((np.arange(D) - D / 2 + 0.5) * dDet), rtol=0, atol=1e-12)
            np.testing.assert_allclose(pw, s2r * dDet * np.cos(gam), rtol=1e-12)
        # the reference's pre-weight ignores binshift (fan_ea_kernel.cu:1075): same values with a shift
        o2 = o.copy()
        o2[10] = 0.7 * dDet
        np.testing.assert_array_equal(O.fbp_preweight(o2)[0], pw)


def test_ed_preweight_geometry():
    """fan_ed_kernel.cu:1349-1355: the ED pre-weight is vd * cos of the angle between bin j's centre ray
    (through the midpoint of its edges on the virtual detector at the isocentre) and the central ray.
    The detector point of each edge is recovered by intersecting the edge ray (source -> mapped axis
    edge) with the detector line, not from the (j - D/2 + 1/2) vd formula."""
    o = fan(0, V=10, D=21, W=10, H=10, dImg=0.1, dDet=0.13, Ang0=0.2, s2r=3.0, d2r=2.0, bs=0.0)
    g = O.parse(o)
    pw, vd = O.fbp_preweight(o)
    assert vd == pytest.approx(0.13 * 3.0 / 5.0, rel=1e-15)
    for v in range(10):
        xmode, rev, e, _ = O.view(o, v)
        th = v * g["dAng"] + g["Ang0"]
        S = np.array([-math.sin(th), math.cos(th)]) * 3.0
        u_dir = np.array([math.cos(th), math.sin(th)])
        P = np.stack([e, np.zeros_like(e)], 1) if xmode else np.stack([np.zeros_like(e), e], 1)
        # S + lam (P - S) on the detector line {u * u_dir} (perpendicular to S): (S + lam (P - S)) . S = 0
        lam = np.array([(S @ S) / ((S - p) @ S) for p in P])
        u = np.array([(S + l_ * (p - S)) @ u_dir for l_, p in zip(lam, P)])
        mid = 0.5 * (u[1:] + u[:-1])
        mid = mid[::-1] if rev else mid
        pc = mid[:, None] * u_dir[None, :] - S[None, :]
        cosang = (pc @ (-S)) / (np.linalg.norm(pc, axis=1) * np.linalg.norm(S))
        np.testing.assert_allclose(pw, vd * cosang, rtol=1e-12)


@pytest.mark.parametrize("st", [0, 1, 2])
def test_fbp_uniform_disk(st):
    """FBP of the projections of a uniform disk recovers the disk value in its interior.

    Fan-EA (fan_ea_kernel.cu:1059-1107) uses the plain RL taps without the equal-angle factor
    (gamma / sin gamma)^2, so the fan is kept narrow (|gamma| <= 0.15 rad: the missing factor is at
    most 1.0075) while still covering the whole image."""
    n = 64
    if st == 2:
        o = para(V=180, D=96, W=n, H=n, dImg=1 / 32, dDet=1 / 32, Ang0=0.0, dAng=math.pi / 180, bs=0.0)
    elif st == 1:
        o = fan(1, V=360, D=128, W=n, H=n, dImg=1 / 32, dDet=0.3 / 128, Ang0=0.0, dAng=2 * math.pi / 360,
                s2r=10.0, d2r=5.0, bs=0.0)
    else:
        o = fan(0, V=360, D=128, W=n, H=n, dImg=1 / 32, dDet=1 / 32, Ang0=0.0, dAng=2 * math.pi / 360,
                s2r=6.0, d2r=3.0, bs=0.0)
    c = (np.arange(n) - n / 2 + 0.5) / (n / 2)
    X, Y = np.meshgrid(c, c)
    mu = 0.7
    disk = np.where(X ** 2 + Y ** 2 <= 0.6 ** 2, mu, 0.0)[None]
    y = O.apply("projection", o, disk)
    rec = O.apply("fbp", o, y)[0]
    inner = X ** 2 + Y ** 2 <= 0.4 ** 2
    scale = 2.0 if st == 2 else 1.0  # fbp scales by dAng/2 for any coverage (para over pi is 2x low)
    val = rec[inner].mean() * scale
    assert abs(val - mu) < 0.02 * mu, val
    outside = X ** 2 + Y ** 2 >= 0.8 ** 2
    assert np.abs(rec[outside]).mean() * scale < 0.02 * mu


def test_parse_options_quirks():
    # 9-entry parallel vector: binshift = options[8] = scan_type = 2 (para_kernel.cu:695, ctlib.cpp:18)
    g = O.parse([30, 21, 16, 16, 0.1, 0.1, 0.0, 0.1, 2])
    assert g["scan_type"] == 2 and g["binshift"] == 2.0
    g = O.parse([30, 21, 16, 16, 0.1, 0.1, 0.0, 0.1, 0.0, 2])
    assert g["binshift"] == 0.0
    with pytest.raises(O.OracleError):
        O.parse([30, 21, 16, 16, 0.1, 0.1, 0.0, 0.1, 0.0, 3])
    with pytest.raises(O.OracleError):
        O.parse([30, 21, 16, 16, 0.1, 0.1, 0.0, 0.1, 5.0, 0])   # fan needs 11+ entries


def test_phantom_generator():
    a = phantom.ellipse_phantom(64, 64, seed=3)
    b = phantom.ellipse_phantom(64, 64, seed=3)
    c = phantom.ellipse_phantom(64, 64, seed=4)
    assert a.dtype == np.float32 and np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() <= 10
    n = phantom.poisson_sinogram(np.full((4, 5), 0.3), 1e5, seed=1)
    assert np.all(np.abs(n - 0.3) < 0.05)
