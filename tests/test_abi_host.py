"""Host-side checks of libctlib_amd.so: it loads, exports every symbol the
public header declares, parses options like the reference, and its FP64
geometry plan - run through a numpy model of the kernels' arithmetic - agrees
with the FP64 oracle.  No GPU is touched."""
import math
import os
import re

import numpy as np
import pytest

from ctlib_amd import abi, phantom
from oracle import oracle as O

import kernel_model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ctlib_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ctlib_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    declared = header_functions()
    assert len(declared) >= 10
    lib = abi.load()
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(abi.EXPORTS) == declared


def test_status_strings():
    lib = abi.load()
    for code in range(8):
        assert lib.ctlib_status_string(code)


@pytest.mark.parametrize("opts,code", [
    ([736, 736, 512, 512, 0.0066, 0.0072, 0, 0.01, 5.95, 4.906, 0, 3], 2),     # bad scan type
    ([736, 736, 512, 512, 0.0066, 0.0072, 0, 0.01, 5.95, 4.906, 0, 0.5], 2),
    ([736, 736, 512, 512, 0.0066, 0.0072, 0, 0.01, 0, 0], 1),                 # fan, too short
    ([736, 736.5, 512, 512, 0.0066, 0.0072, 0, 0.01, 5.95, 4.906, 0, 0], 1),   # non-integral size
    ([736, 736, 512, 512, -0.0066, 0.0072, 0, 0.01, 5.95, 4.906, 0, 0], 1),    # negative pixel
    ([736, 736, 512, 512, 0.0066, 0.0072, 0, 0.01, 2], 2 - 2),                  # 9-entry para: valid
])
def test_parse_errors(opts, code):
    if code == 0:
        g = abi.parse_options(opts)
        assert g.scan_type == 2 and g.binshift == 2.0   # README 9-entry vector: binshift = scan_type
        return
    with pytest.raises(abi.CtlibError) as e:
        abi.parse_options(opts)
    assert e.value.status == code


def test_plan_rejects_source_inside_image():
    opts = phantom.fan_options(16, 16, 64, 64, 0.1, 0.1, 2.0, 1.0, 0)  # source at r=2 inside a 6.4 wide image
    with pytest.raises(abi.CtlibError) as e:
        abi.plan_export(opts, "edges")
    assert e.value.status == 7


@pytest.mark.parametrize("st", [0, 1, 2])
def test_plan_axis_decisions_match_oracle(st):
    """Axis and reversal flags per view, including views on pi/4 ties."""
    if st == 2:
        o = phantom.para_options(8, 19, 10, 10, 0.1, 0.1, Ang0=0.0, dAng=math.pi / 4)
    else:
        o = phantom.fan_options(8, 21, 10, 10, 0.1, 0.03 if st == 1 else 0.13, 3.0, 2.0, st, Ang0=0.0,
                                dAng=math.pi / 4)
    vp = abi.plan_export(o, "views")
    edges = abi.plan_export(o, "edges").reshape(8, -1)
    for v in range(8):
        xm, rev, e, _ = O.view(o.astype(np.float64), v)
        assert bool(vp[v]["xmode"]) == xm and bool(vp[v]["rev"]) == rev
        np.testing.assert_allclose(edges[v], e, rtol=1e-14, atol=1e-15)


def test_tie_views_of_measured_configs():
    """Views within 1e-7 of the pi/4 axis threshold (config 0 views 45, 135; config 4 views 180, 540)
    take the axis the FP64 oracle takes."""
    for cfg, views in ((0, (45, 135)), (4, (180, 540))):
        o = phantom.CONFIGS[cfg]["options"]
        vp = abi.plan_export(o, "views")
        for v in views:
            xm, rev, _, _ = O.view(o.astype(np.float64), v)
            assert bool(vp[v]["xmode"]) == xm and bool(vp[v]["rev"]) == rev


def small(st, **kw):
    V, D, W, H = kw.get("V", 24), kw.get("D", 17), kw.get("W", 12), kw.get("H", 10)
    if st == 2:
        return phantom.para_options(V, D, W, H, 0.1, 0.13, Ang0=kw.get("Ang0", 0.1), dAng=2 * math.pi / V,
                                    binshift=kw.get("bs", 0.02))
    return phantom.fan_options(V, D, W, H, 0.1, 0.03 if st == 1 else 0.13, 3.0, 2.0, st, Ang0=kw.get("Ang0", 0.1),
                               binshift=kw.get("bs", 0.02))


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("op", ["projection", "projection_t", "backprojection", "backprojection_t",
                                "backprojection_t_true", "fbp"])
def test_kernel_arithmetic_model_matches_oracle(st, op):
    """The kernels' cumulative-integral arithmetic on the exported plan equals the oracle's merge."""
    o = small(st, V=30, D=19, W=14, H=11, Ang0=0.25, bs=-0.03)
    P = M.Plan(o)
    g = P.g
    rng = np.random.default_rng(12)
    if op in ("projection", "backprojection_t", "backprojection_t_true"):
        x = rng.random((2, g["height"], g["width"]))
    else:
        x = rng.random((2, g["views"], g["dets"]))
    got = M.apply(P, op, x)
    ref = O.apply(op, o.astype(np.float64), x)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    tol = 1e-13 if op in ("projection", "projection_t") or st == 2 else 1e-7
    assert err < tol, (op, st, err)


def test_kernel_model_parallel_rows_at_full_config0():
    """Config 0 (parallel 256^2, 180 views, 367 bins): model vs oracle on a few views' worth of data."""
    o = phantom.CONFIGS[0]["options"]
    P = M.Plan(o)
    x = phantom.phantom_batch(1, 256, 256, config=0)[:, 0].astype(np.float64)
    got = M.apply(P, "projection", x)
    ref = O.apply("projection", o.astype(np.float64), x)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("st", [0, 1, 2])
def test_composed_lookup_matches_view_lookup(st):
    """Plan table adj_lk - each view's bin lookup composed with each line's pixel-edge map, which the
    fp32 bin->pixel kernels evaluate at pixel edge p as (c0 p/64 + c1) / (c2 p/64 + c3) - gives the
    lookup argument of the view's own lookup at a + s p; for fan-ED and parallel beam that lookup
    puts the bin edges e_k of the (oracle-checked) edge table at k + 1."""
    o = small(st, V=30, D=19, W=14, H=11, Ang0=0.25, bs=-0.03)
    g = abi.parse_options(o)
    V, D = g.views, g.dets
    nmax = max(g.width, g.height)
    vp = abi.plan_export(o, "views")
    lines = abi.plan_export(o, "lines").reshape(V, nmax)
    lk = abi.plan_export(o, "adj_lk").reshape(V, nmax, 4).astype(np.float64)
    edges = abi.plan_export(o, "edges").reshape(V, D + 1)
    for v in range(V):
        c = vp[v]["lk"].astype(np.float64)
        view_lk = lambda e: (c[0] * e + c[1]) / (c[2] * e + c[3])
        if st != 1:
            np.testing.assert_allclose(view_lk(edges[v]), np.arange(1, D + 2), rtol=0, atol=2e-4)
        p = np.arange(vp[v]["npix"] + 1, dtype=np.float64)
        for l in range(vp[v]["nlines"]):
            want = view_lk(lines[v, l]["a"] + lines[v, l]["s"] * p)
            k = lk[v, l]
            got = (k[0] * p / 64 + k[1]) / (k[2] * p / 64 + k[3])
            np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * max(1.0, np.max(np.abs(want))))


@pytest.mark.parametrize("st", [0, 1, 2])
def test_prescaled_pixel_weights(st):
    """Plan table adj_wk - each line's pixel weight rewritten as {64 c w1, c w1, c w0, c^2 w2} with
    c = sqrt(s), which the fp32 weighted bin->pixel kernel evaluates at pixel p = j + 64 k as
    1 / (u^2 + c^2 w2), u = k (64 c w1) + j c w1 + c w0 (fp32 FMAs) - equals the weight the oracle
    and the fp64 path use, (1/s) / ((w0 + w1 p)^2 + w2) from the line's fp64 coefficients, to fp32
    rounding; for parallel beam it is 1/s; for fan-ED it is 1/(s u^2) with u the distance ratio
    1 - <source, pixel centre>/R^2 (w2 = 0)."""
    # 150 x 130 pixels of 0.01: lines over two 64-pixel groups
    if st == 2:
        o = phantom.para_options(30, 19, 150, 130, 0.01, 0.13, Ang0=0.25, dAng=2 * math.pi / 30, binshift=-0.03)
    else:
        o = phantom.fan_options(30, 19, 150, 130, 0.01, 0.03 if st == 1 else 0.13, 3.0, 2.0, st, Ang0=0.25,
                                binshift=-0.03)
    g = abi.parse_options(o)
    V = g.views
    nmax = max(g.width, g.height)
    vp = abi.plan_export(o, "views")
    lines = abi.plan_export(o, "lines").reshape(V, nmax)
    wk = abi.plan_export(o, "adj_wk").reshape(V, nmax, 4)
    for v in range(V):
        n = int(vp[v]["npix"])
        p = np.arange(n)
        j, k = (p % 64).astype(np.float32), (p // 64).astype(np.float32)
        for l in range(int(vp[v]["nlines"])):
            c = lines[v, l]
            want = c["inv_s"] / ((c["w0"] + c["w1"] * p) ** 2 + c["w2"])
            q = wk[v, l]
            u0 = (q[1] * j + q[2]).astype(np.float32)
            u = (k * q[0] + u0).astype(np.float32)
            got = 1.0 / (u.astype(np.float64) ** 2 + np.float64(q[3]))
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)
            if st == 2:
                np.testing.assert_allclose(want, c["inv_s"], rtol=1e-15)


@pytest.mark.parametrize("dDet, expect_poly", [(0.03, True), (0.0011, True), (0.0005, True), (0.0016, False)])
def test_fan_ea_atan_polynomial(dDet, expect_poly):
    """The fan-EA lookup replaces atanf by an odd polynomial on |t| <= T: T must cover every bin edge's
    fan angle (plus one bin), and the fp32 coefficients must stay within 5e-8 rad of atan there;
    wider fans keep atanf."""
    D = 888 if dDet < 0.01 else 17
    o = phantom.fan_options(64, D, 32, 32, 0.05 if dDet >= 0.01 else 0.0066, dDet, 5.95, 4.906, 1, binshift=0.0003)
    ea = abi.plan_export(o, "ea_atan")
    c, T, use = ea[:6].astype(np.float64), float(ea[6]), bool(ea[7])
    assert use == expect_poly
    if not use:
        return
    gmax = max(abs(-D / 2 * dDet + 0.0003), abs(D / 2 * dDet + 0.0003))
    assert T >= np.tan(gmax)
    t = np.linspace(0.0, T, 20001)
    p = t * np.polyval(c[::-1], t * t)
    assert np.max(np.abs(p - np.arctan(t))) <= 5e-8
    # monotone on [0, T], so clamped rays beyond T stay beyond the last edge
    assert np.all(np.diff(p) > 0)


def test_laplacian_argument_checks_without_gpu():
    """ctlib_laplacian validates its arguments before any device work."""
    import ctypes
    lib = abi.load()
    lib.ctlib_laplacian.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_void_p]
    lib.ctlib_laplacian_workspace_bytes.restype = ctypes.c_size_t
    lib.ctlib_laplacian_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    fake = ctypes.c_void_p(0x1000)
    # k > n, k < 1, k above the supported 4096, bad dtype, missing workspace
    assert lib.ctlib_laplacian(fake, 10, 4, 11, 0, fake, fake, fake, 1 << 20, None) == 1
    assert lib.ctlib_laplacian(fake, 10, 4, 0, 0, fake, fake, fake, 1 << 20, None) == 1
    assert lib.ctlib_laplacian(fake, 10000, 4, 4097, 0, fake, fake, fake, 1 << 20, None) == 1
    assert lib.ctlib_laplacian(fake, 10, 4, 4, 7, fake, fake, fake, 1 << 20, None) == 5
    assert lib.ctlib_laplacian(fake, 10, 4, 4, 0, fake, fake, None, 0, None) == 6
    assert lib.ctlib_laplacian_workspace_bytes(1000, 16, 8, 0) >= 1000 * 4 + 1000 * 8 * 4
    # any dimension: the general path needs a chunk of distance rows (here all 1000 rows)
    assert lib.ctlib_laplacian_workspace_bytes(1000, 300, 8, 0) >= 1000 * 1000 * 4
    assert lib.ctlib_laplacian_workspace_bytes(1000, 16, 4097, 0) == 0
