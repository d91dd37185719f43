"""GPU parity of every operator against the FP64 oracle (oracle/ct_oracle.cpp).

Bars (fp32 data): relative L2 <= 1e-5 for every operator; adjoint dot-product
identities <= 1e-6 relative.  fp64 data: relative L2 <= 1e-7 (the pixel
weights of backprojection / backprojection_t use a fp32 reciprocal).
"""
import math

import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import phantom
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5
F64_TOL = 1e-7
ADJ_TOL = 1e-6

IMG_IN = ("projection", "backprojection_t", "backprojection_t_true", "fbp_t")
OPS = ("projection", "projection_t", "backprojection", "backprojection_t", "backprojection_t_true", "fbp", "fbp_t")


def small_options(st, V=24, D=17, W=12, H=10, Ang0=0.1, bs=0.02):
    if st == 2:
        return phantom.para_options(V, D, W, H, 0.1, 0.13, Ang0=Ang0, dAng=2 * math.pi / V, binshift=bs)
    return phantom.fan_options(V, D, W, H, 0.1, 0.03 if st == 1 else 0.13, 3.0, 2.0, st, Ang0=Ang0, binshift=bs)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def gpu_op(op, x, options, dev):
    t = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    o = torch.from_numpy(np.asarray(options))
    return getattr(ctlib_amd, op)(t, o).cpu().numpy()


def oracle_op(op, x, options):
    if op == "fbp_t":
        # adjoint of fbp = preweight * ramp(dAng/2 * B_true^T x)
        z = O.apply("backprojection_t_true", options, x[:, 0]) * (float(options[7]) / 2)
        pw, delta = O.fbp_preweight(options)
        h = O.rl_taps(z.shape[-1], delta)
        D = z.shape[-1]
        T = h[(np.arange(D)[:, None] - np.arange(D)[None, :]) + D - 1]
        return ((z @ T) * pw)[:, None]
    return O.apply(op, options, x[:, 0])[:, None]


def inputs(op, options, B, rng, dtype=np.float32):
    g = O.parse(options)
    shape = (B, 1, g["height"], g["width"]) if op in IMG_IN else (B, 1, g["views"], g["dets"])
    return rng.random(shape).astype(dtype)


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("op", OPS)
def test_small_geometry_f32(dev, st, op):
    rng = np.random.default_rng(100 + st)
    o = small_options(st)
    x = inputs(op, o, 3, rng)
    print(f"[case] small st={st} op={op} f32", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    assert got.dtype == np.float32 and got.shape == ref.shape
    assert rel_l2(got, ref) <= F32_TOL, (op, st, rel_l2(got, ref))


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("op", OPS)
def test_small_geometry_f64(dev, st, op):
    rng = np.random.default_rng(200 + st)
    o = small_options(st, V=40, D=23, W=16, H=20, Ang0=0.3, bs=-0.05)
    x = inputs(op, o, 2, rng, np.float64)
    print(f"[case] small st={st} op={op} f64", flush=True)
    got = gpu_op(op, x, o.astype(np.float64), dev)
    ref = oracle_op(op, x, o.astype(np.float64))
    assert got.dtype == np.float64
    assert rel_l2(got, ref) <= F64_TOL, (op, st, rel_l2(got, ref))


@pytest.mark.parametrize("op", OPS)
def test_config0_parallel(dev, op):
    """Parallel beam, 256^2, 180 views over pi, 367 bins (ellipse phantom input)."""
    o = phantom.CONFIGS[0]["options"]
    rng = np.random.default_rng(7)
    if op in IMG_IN:
        x = phantom.phantom_batch(1, 256, 256, config=0)
    else:
        clean = O.apply("projection", o, phantom.phantom_batch(1, 256, 256, config=0)[:, 0].astype(np.float64))
        x = (clean[:, None] + 0.01 * rng.random(clean[:, None].shape)).astype(np.float32)
    print(f"[case] config0 op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    err = rel_l2(got, ref)
    print(f"  rel_l2={err:.3e} max_abs={np.abs(got - ref).max():.3e}")
    assert err <= F32_TOL


@pytest.mark.parametrize("cfg", [1, 2, 4])
@pytest.mark.parametrize("op", ["projection", "projection_t", "backprojection", "backprojection_t", "fbp"])
def test_full_fan_geometry(dev, cfg, op):
    """Fan-ED 512^2 736x736, fan-EA 512^2 1024x888 and fan-ED 1024^2 1440x1536 geometries (one batch slice)."""
    o = phantom.CONFIGS[cfg]["options"]
    g = O.parse(o)
    x_img = phantom.phantom_batch(1, g["height"], g["width"], config=cfg, scale=0.02)
    if op in IMG_IN:
        x = x_img
    else:
        x = O.apply("projection", o, x_img[:, 0].astype(np.float64))[:, None].astype(np.float32)
    print(f"[case] config{cfg} op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    err = rel_l2(got, ref)
    print(f"  rel_l2={err:.3e} max_abs={np.abs(got - ref).max():.3e}")
    assert err <= F32_TOL


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("wh", [(256, 128), (128, 256)])
@pytest.mark.parametrize("op", ["projection", "backprojection", "backprojection_t", "backprojection_t_true", "fbp_t"])
def test_rectangular_images(dev, st, wh, op):
    """Rows and columns of different lengths, one of them a whole number of 64-pixel lane groups
    (the weighted line->bin kernel takes its bounds-checked form unless every line is full)."""
    W, H = wh
    if st == 2:
        o = phantom.para_options(48, 97, W, H, 0.01, 0.03, Ang0=0.2, dAng=math.pi / 48, binshift=0.01)
    else:
        o = phantom.fan_options(48, 97, W, H, 0.01, 0.012 if st == 1 else 0.06, 3.0, 2.0, st, Ang0=0.2,
                                binshift=0.01)
    rng = np.random.default_rng(300 + st)
    x = inputs(op, o, 2, rng)
    print(f"[case] rectangular {W}x{H} st={st} op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    assert rel_l2(got, ref) <= F32_TOL, (op, st, wh, rel_l2(got, ref))


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("op", ["backprojection_t", "backprojection_t_true", "fbp_t", "projection"])
def test_lines_longer_than_1024(dev, st, op):
    """Rows of 1152 pixels: the weighted line->bin operators take the shared-table kernel with one
    table per view (the wave-private-table kernel covers lines of up to 1024 pixels)."""
    if st == 2:
        o = phantom.para_options(12, 97, 1152, 48, 0.002, 0.03, Ang0=0.2, dAng=math.pi / 12, binshift=0.01)
    else:
        o = phantom.fan_options(12, 97, 1152, 48, 0.002, 0.012 if st == 1 else 0.06, 3.0, 2.0, st, Ang0=0.2,
                                binshift=0.01)
    rng = np.random.default_rng(500 + st)
    x = inputs(op, o, 2, rng)
    print(f"[case] 1152-pixel rows st={st} op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    assert rel_l2(got, ref) <= F32_TOL, (op, st, rel_l2(got, ref))


@pytest.mark.parametrize("n", [384, 400, 768])
@pytest.mark.parametrize("st", [0, 1])
@pytest.mark.parametrize("op", ["projection_t", "backprojection"])
def test_square_line_lengths(dev, n, st, op):
    """384- and 768-pixel lines (a whole number of 64-pixel groups: the last pixel edge is evaluated
    packed) and 400-pixel lines (a partly filled last group)."""
    o = phantom.fan_options(12, 97, n, n, 2.3 / n, 0.012 if st == 1 else 0.06, 3.0, 2.0, st, Ang0=0.2,
                            binshift=0.01)
    rng = np.random.default_rng(600 + n + st)
    x = inputs(op, o, 1, rng)
    print(f"[case] {n}^2 st={st} op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    assert rel_l2(got, ref) <= F32_TOL, (op, st, n, rel_l2(got, ref))


def test_line_tables_setting(dev):
    """ctlib_plan_set_line_tables: 0 (prebuilt tables, default) or 1 (built per block); anything else
    is refused.  The prebuilt tables live in the workspace: with the same launch split, projection's
    workspace grows by B * (H + W) * (256 M + 2) * 16 bytes (M = 2 for 512-pixel lines), and the
    per-block setting needs none of it."""
    from ctlib_amd import _C
    o = phantom.CONFIGS[1]["options"]
    ot = torch.from_numpy(o)
    d = dev.index or 0
    g = O.parse(o)
    B = 4
    with pytest.raises(Exception):
        _C.set_plan_line_tables(ot, d, 2)
    try:
        _C.set_plan_chunks(ot, d, 1, 1)  # same (unsplit) launches for both settings
        _C.set_plan_line_tables(ot, d, 1)
        per_block = _C.workspace_bytes(ot, _C.OP_PRJ, B, False, d)
        _C.set_plan_line_tables(ot, d, 0)
        prebuilt = _C.workspace_bytes(ot, _C.OP_PRJ, B, False, d)
    finally:
        _C.set_plan_line_tables(ot, d, 0)
        _C.set_plan_chunks(ot, d, 0, 0)
    tables = B * (g["width"] + g["height"]) * (256 * 2 + 2) * 16
    assert prebuilt - per_block >= tables
    assert prebuilt - per_block < tables + 4096


@pytest.mark.parametrize("per_block", [0, 1])
@pytest.mark.parametrize("wh", [(256, 256), (300, 40), (40, 640), (1024, 64), (1800, 48), (48, 2048)])
@pytest.mark.parametrize("st", [0, 2])
def test_forward_line_lengths(dev, st, wh, per_block):
    """Projection and the residual step over every pixels-per-thread class of the line->bin
    kernel (lines of 256, 512, 1024 and 2048 pixels at most; whole 256-pixel multiples or not;
    the two sweep axes of one slice of different lengths), with line tables prebuilt in the
    workspace (per_block 0) or built by each block (1)."""
    from ctlib_amd import _C
    W, H = wh
    if st == 2:
        o = phantom.para_options(16, 97, W, H, 2.0 / max(W, H), 0.03, Ang0=0.2, dAng=math.pi / 16, binshift=0.01)
    else:
        o = phantom.fan_options(16, 97, W, H, 2.0 / max(W, H), 0.06, 3.0, 2.0, st, Ang0=0.2, binshift=0.01)
    rng = np.random.default_rng(700 + W + H + st)
    x = rng.random((2, 1, H, W)).astype(np.float32)
    mm = rng.random((2, 1, 16, 97)).astype(np.float32)
    ot = torch.from_numpy(o)
    print(f"[case] forward line lengths {W}x{H} st={st} per_block={per_block}", flush=True)
    try:
        _C.set_plan_line_tables(ot, dev.index or 0, per_block)
        xt, mt = torch.from_numpy(x).to(dev), torch.from_numpy(mm).to(dev)
        y = ctlib_amd.projection(xt, ot).cpu().numpy()
        yr, gr = ctlib_amd.residual_step(xt, mt, ot)
    finally:
        _C.set_plan_line_tables(ot, dev.index or 0, 0)
    o64 = o.astype(np.float64)
    y_ref = O.apply("projection", o64, x[:, 0].astype(np.float64))
    g_ref = O.apply("projection_t", o64, y_ref - mm[:, 0].astype(np.float64))
    assert rel_l2(y[:, 0], y_ref) <= F32_TOL
    assert rel_l2(yr.cpu().numpy()[:, 0], y_ref) <= F32_TOL
    assert rel_l2(gr.cpu().numpy()[:, 0], g_ref) <= F32_TOL


@pytest.mark.parametrize("D", [512, 1024, 1536])
@pytest.mark.parametrize("st", [0, 2])
@pytest.mark.parametrize("op", ["projection", "backprojection_t"])
def test_detector_multiple_of_64(dev, D, st, op):
    """Detectors of 64*k bins (their last edge is evaluated apart from the 64-edge groups)."""
    if st == 2:
        o = phantom.para_options(20, D, 64, 64, 0.02, 1.6 / D, Ang0=0.3, dAng=math.pi / 20, binshift=0.001)
    else:
        o = phantom.fan_options(20, D, 64, 64, 0.02, 2.6 / D, 3.0, 2.0, st, Ang0=0.3, binshift=0.001)
    rng = np.random.default_rng(400 + D + st)
    x = inputs(op, o, 2, rng)
    print(f"[case] D={D} st={st} op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x.astype(np.float64), o.astype(np.float64))
    assert rel_l2(got, ref) <= F32_TOL, (op, st, D, rel_l2(got, ref))


@pytest.mark.parametrize("st", [0, 1, 2])
def test_adjoint_dot(dev, st):
    o = small_options(st, V=64, D=41, W=32, H=24)
    g = O.parse(o)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.random((2, 1, g["height"], g["width"])).astype(np.float32)).to(dev)
    y = torch.from_numpy(rng.random((2, 1, g["views"], g["dets"])).astype(np.float32)).to(dev)
    ot = torch.from_numpy(o)
    print(f"[case] adjoint st={st}", flush=True)
    pairs = [("projection", "projection_t"), ("backprojection_t_true", "backprojection"), ("fbp_t", "fbp")]
    for fwd, adj in pairs:
        lhs = float((getattr(ctlib_amd, fwd)(x, ot).double() * y.double()).sum())
        rhs = float((x.double() * getattr(ctlib_amd, adj)(y, ot).double()).sum())
        assert abs(lhs - rhs) <= ADJ_TOL * max(abs(lhs), abs(rhs)), (fwd, adj, lhs, rhs)


def test_deterministic(dev):
    o = phantom.CONFIGS[1]["options"]
    x = torch.from_numpy(phantom.phantom_batch(2, 512, 512, config=1, scale=0.02)).to(dev)
    ot = torch.from_numpy(o)
    print("[case] determinism", flush=True)
    y1 = ctlib_amd.projection(x, ot)
    y2 = ctlib_amd.projection(x, ot)
    assert torch.equal(y1, y2)
    a1 = ctlib_amd.projection_t(y1, ot)
    a2 = ctlib_amd.projection_t(y1, ot)
    assert torch.equal(a1, a2)


def test_device_options_and_channels(dev):
    o = small_options(0)
    g = O.parse(o)
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.random((2, 3, g["height"], g["width"])).astype(np.float32)).to(dev)
    print("[case] device options / channel 0", flush=True)
    host = ctlib_amd.projection(x[:, :1].contiguous(), torch.from_numpy(o))
    devo = ctlib_amd.projection(x, torch.from_numpy(o).to(dev))   # options on the device, 3 channels
    assert devo.shape == host.shape
    assert torch.equal(host, devo)


def test_para_nine_entry_options_binshift_quirk(dev):
    """The README's 9-entry parallel vector reads binshift = options[8] = scan_type = 2."""
    o10 = phantom.para_options(30, 21, 16, 16, 0.1, 0.1, binshift=2.0)
    o9 = np.concatenate([o10[:8], [2.0]]).astype(np.float32)
    rng = np.random.default_rng(9)
    x = rng.random((1, 1, 16, 16)).astype(np.float32)
    print("[case] 9-entry para options", flush=True)
    a = gpu_op("projection", x, o9, dev)
    b = gpu_op("projection", x, o10, dev)
    np.testing.assert_array_equal(a, b)
    ref = oracle_op("projection", x.astype(np.float64), o9.astype(np.float64))
    assert rel_l2(a, ref) <= F32_TOL


def test_errors(dev):
    o = torch.from_numpy(small_options(0))
    x = torch.zeros(1, 1, 10, 12, device=dev)
    print("[case] errors", flush=True)
    with pytest.raises(RuntimeError):
        ctlib_amd.projection(torch.zeros(1, 1, 10, 12), o)           # CPU tensor
    with pytest.raises(RuntimeError):
        ctlib_amd.projection(torch.zeros(1, 1, 12, 10, device=dev).transpose(2, 3), o)  # non-contiguous
    with pytest.raises(ValueError):
        ctlib_amd.projection(torch.zeros(1, 1, 11, 12, device=dev), o)  # shape mismatch
    bad = o.clone()
    bad[-1] = 7
    with pytest.raises(ValueError):
        ctlib_amd.projection(x, bad)
    with pytest.raises(TypeError):
        ctlib_amd.projection(x.half(), o)
    assert ctlib_amd.projection(torch.zeros(0, 1, 10, 12, device=dev), o).shape == (0, 1, 24, 17)


def test_residual_step(dev):
    o = small_options(1, V=48, D=31, W=20, H=20)
    g = O.parse(o)
    rng = np.random.default_rng(11)
    x = torch.from_numpy(rng.random((2, 1, 20, 20)).astype(np.float32)).to(dev)
    m = torch.from_numpy(rng.random((2, 1, g["views"], g["dets"])).astype(np.float32)).to(dev)
    ot = torch.from_numpy(o)
    print("[case] residual_step", flush=True)
    y, grad = ctlib_amd.residual_step(x, m, ot)
    y_ref = ctlib_amd.projection(x, ot)
    g_ref = ctlib_amd.projection_t(y_ref - m, ot)
    assert torch.equal(y, y_ref)
    assert rel_l2(grad.cpu().numpy(), g_ref.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("per_block", [0, 1])
@pytest.mark.parametrize("G", [4, 8])
@pytest.mark.parametrize("case", ["ed_small", "ea_odd", "para_ties", "config1"])
def test_forward_launch_shapes(dev, case, G, per_block):
    """Both line->bin launch shapes the autotuner may pick (4 or 8 views per block), over line
    tables prebuilt in the workspace (per_block 0) or built by each block (1), give the
    oracle's projection and residual step."""
    from ctlib_amd import _C
    if case == "config1":
        o = phantom.CONFIGS[1]["options"]
    elif case == "ed_small":
        o = small_options(0, V=40, D=23, W=16, H=20)
    elif case == "ea_odd":
        o = small_options(1, V=33, D=31, W=13, H=11, Ang0=0.7)   # odd sizes, a partial view group
    else:
        o = phantom.para_options(8, 19, 10, 10, 0.1, 0.13, Ang0=0.0, dAng=math.pi / 4, binshift=0.0)
    g = O.parse(o)
    rng = np.random.default_rng(17)
    x = rng.random((2, 1, g["height"], g["width"])).astype(np.float32)
    mm = rng.random((2, 1, g["views"], g["dets"])).astype(np.float32)
    ot = torch.from_numpy(o)
    print(f"[case] forward views per block {G} line tables per block {per_block} {case}", flush=True)
    try:
        _C.set_plan_tuning(ot, dev.index or 0, G)
        _C.set_plan_line_tables(ot, dev.index or 0, per_block)
        xt, mt = torch.from_numpy(x).to(dev), torch.from_numpy(mm).to(dev)
        y = ctlib_amd.projection(xt, ot).cpu().numpy()
        yr, gr = ctlib_amd.residual_step(xt, mt, ot)
    finally:
        _C.set_plan_tuning(ot, dev.index or 0, 0)
        _C.set_plan_line_tables(ot, dev.index or 0, 0)
    o64 = o.astype(np.float64)
    y_ref = O.apply("projection", o64, x[:, 0].astype(np.float64))
    g_ref = O.apply("projection_t", o64, y_ref - mm[:, 0].astype(np.float64))
    assert rel_l2(y[:, 0], y_ref) <= F32_TOL
    assert rel_l2(yr.cpu().numpy()[:, 0], y_ref) <= F32_TOL
    assert rel_l2(gr.cpu().numpy()[:, 0], g_ref) <= F32_TOL


_CHUNK_REF = {}


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("nch", [0, 1, 2, 3, 7, 64])
@pytest.mark.parametrize("case", ["ed_small", "ea_odd", "para_ties", "config1"])
def test_line_and_view_chunks(dev, case, nch, dt):
    """Launches split into line chunks (line->bin) and view chunks (bin->pixel), with the
    partial sinograms / images summed afterwards, give the oracle's result for every
    operator and the residual step, at batch 1 and 2 (64 is capped at the lines / views
    of the smaller sweep axis; 0 is the automatic choice), for fp32 and fp64 data."""
    from ctlib_amd import _C
    npdt, tol = (np.float32, F32_TOL) if dt == "f32" else (np.float64, F64_TOL)
    if case == "config1":
        o = phantom.CONFIGS[1]["options"]
    elif case == "ed_small":
        o = small_options(0, V=40, D=23, W=16, H=20)
    elif case == "ea_odd":
        o = small_options(1, V=33, D=31, W=13, H=11, Ang0=0.7)
    else:
        o = phantom.para_options(8, 19, 10, 10, 0.1, 0.13, Ang0=0.0, dAng=math.pi / 4, binshift=0.0)
    rng = np.random.default_rng(23)
    ot = torch.from_numpy(o)
    o64 = o.astype(np.float64)
    print(f"[case] line and view chunks {nch} {case} {dt}", flush=True)
    for B in (1, 2):
        ins = {op: inputs(op, o, B, rng, npdt) for op in OPS}
        got = {}
        try:
            _C.set_plan_tuning(ot, dev.index or 0, 0)
            _C.set_plan_chunks(ot, dev.index or 0, nch, nch)
            for op in OPS:
                got[op] = gpu_op(op, ins[op], o, dev)
            x, y = ins["projection"], ins["projection_t"]
            _, gr = ctlib_amd.residual_step(torch.from_numpy(x).to(dev), torch.from_numpy(y).to(dev), ot)
            got["residual"] = gr.cpu().numpy()
        finally:
            _C.set_plan_chunks(ot, dev.index or 0, 0, 0)
        # the inputs depend only on (case, dt, B), so the oracle's results are shared by every nch
        for op in OPS:
            key = (case, dt, B, op)
            if key not in _CHUNK_REF:
                _CHUNK_REF[key] = oracle_op(op, ins[op].astype(np.float64), o64)
            ref = _CHUNK_REF[key]
            assert got[op].dtype == npdt
            assert rel_l2(got[op], ref) <= tol, (op, B, rel_l2(got[op], ref))
        key = (case, dt, B, "residual")
        if key not in _CHUNK_REF:
            ax = O.apply("projection", o64, x[:, 0].astype(np.float64))
            _CHUNK_REF[key] = O.apply("projection_t", o64, ax - y[:, 0].astype(np.float64))
        assert rel_l2(got["residual"][:, 0], _CHUNK_REF[key]) <= tol, ("residual", B)


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("cfg", [0, 1, 2])
@pytest.mark.parametrize("op", ["projection_t", "backprojection", "fbp"])
def test_full_geometry_random(dev, cfg, op, dt):
    """Uniform random sinograms on the full parallel 256^2, fan-ED 512^2 and fan-EA 512^2
    geometries (fbp's ramp makes the output cancel heavily).  fp64 data: the bin->pixel kernels
    take their table index from fp64 lookup coefficients, so the result holds the fp64
    tolerance; fp32 data holds the fp32 one with the fp32 index."""
    o = phantom.CONFIGS[cfg]["options"].astype(np.float64)
    g = O.parse(o)
    rng = np.random.default_rng(300 + cfg)
    y = rng.random((1, 1, g["views"], g["dets"]))
    if dt == "f32":
        y = y.astype(np.float32)
    print(f"[case] {dt} random config{cfg} op={op}", flush=True)
    got = gpu_op(op, y, o, dev)
    ref = oracle_op(op, y.astype(np.float64), o)
    assert got.dtype == y.dtype
    err = rel_l2(got, ref)
    print(f"  rel_l2={err:.3e}")
    assert err <= (F64_TOL if dt == "f64" else F32_TOL), err


@pytest.mark.parametrize("case", ["config0", "bins1536"])
@pytest.mark.parametrize("op", ["projection", "backprojection_t", "backprojection_t_true"])
def test_f64_special_slot_counts(dev, case, op):
    """fp64 data through the forward variants with few or many edge slots: the parallel 256^2
    geometry with 367 bins (6-slot variants, one pixel per thread in the line tables) and a
    1536-bin detector over 1024-pixel lines (the 24-slot forward and the 26-slot weighted one)."""
    if case == "config0":
        o = phantom.CONFIGS[0]["options"].astype(np.float64)
    else:
        o = phantom.fan_options(24, 1536, 1024, 1024, 0.0033, 0.0036, 5.95, 4.906, 0).astype(np.float64)
    g = O.parse(o)
    x = np.random.default_rng(800).random((1, 1, g["height"], g["width"]))
    print(f"[case] fp64 {case} op={op}", flush=True)
    got = gpu_op(op, x, o, dev)
    ref = oracle_op(op, x, o)
    assert got.dtype == np.float64
    err = rel_l2(got, ref)
    print(f"  rel_l2={err:.3e}")
    assert err <= F64_TOL, err
