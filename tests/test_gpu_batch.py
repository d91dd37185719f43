"""Heterogeneous batch: one persistent fused forward + adjoint launch over slices
of different image size, views, bins and scan type, replayed from a HIP graph.
Checked against the per-slice plain path and the FP64 oracle; schedules and
graph replays must not change the result."""
import math

import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import batch, phantom, tune
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_slices():
    """Eight small slices covering every scan type, both sweep axes, odd/even sizes and a non-square image."""
    o = [
        phantom.fan_options(60, 47, 32, 32, 0.05, 0.06, 3.0, 2.0, 0),
        phantom.fan_options(45, 64, 48, 40, 0.04, 0.02, 3.0, 2.0, 1, Ang0=0.3, binshift=0.01),
        phantom.para_options(36, 53, 40, 40, 0.03, 0.03),
        phantom.fan_options(90, 96, 64, 64, 0.03, 0.035, 4.0, 2.5, 0, binshift=-0.02),
        phantom.para_options(7, 33, 24, 24, 0.05, 0.05, Ang0=0.0, dAng=math.pi / 7),
        phantom.fan_options(120, 130, 96, 96, 0.02, 0.012, 5.95, 4.906, 1),
        phantom.fan_options(8, 21, 16, 16, 0.1, 0.12, 3.0, 2.0, 0, Ang0=0.0, dAng=math.pi / 4),   # tie views
        phantom.para_options(64, 100, 80, 72, 0.02, 0.02, binshift=0.005),
    ]
    rng = np.random.default_rng(77)
    xs, ms = [], []
    for oo in o:
        g = O.parse(oo)
        xs.append(rng.random((1, 1, g["height"], g["width"])).astype(np.float32))
        ms.append(rng.random((1, 1, g["views"], g["dets"])).astype(np.float32))
    return o, xs, ms


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_batch_matches_plain_and_oracle(dev):
    o, xs, ms = small_slices()
    print("[case] batch small: build", flush=True)
    hb = batch.HeteroBatch([torch.from_numpy(oo) for oo in o], device=dev, graph=False)
    print(f"[case] batch small: run ({hb.info()['items']} items)", flush=True)
    ys, grads = hb.run([torch.from_numpy(x).to(dev) for x in xs], [torch.from_numpy(m).to(dev) for m in ms])
    hb.check()
    for s, oo in enumerate(o):
        ot = torch.from_numpy(oo)
        xt, mt = torch.from_numpy(xs[s]).to(dev), torch.from_numpy(ms[s]).to(dev)
        y_ref, g_ref = ctlib_amd.residual_step(xt, mt, ot)
        assert rel(ys[s].cpu().numpy(), y_ref.cpu().numpy()) < 1e-6, s
        assert rel(grads[s].cpu().numpy(), g_ref.cpu().numpy()) < 1e-6, s
        o64 = oo.astype(np.float64)
        y_or = O.apply("projection", o64, xs[s][:, 0].astype(np.float64))
        g_or = O.apply("projection_t", o64, y_or - ms[s][:, 0].astype(np.float64))
        assert rel(ys[s].cpu().numpy()[:, 0], y_or) < 1e-5, s
        assert rel(grads[s].cpu().numpy()[:, 0], g_or) < 1e-5, s


def test_batch_schedules_and_graph_replays_identical(dev):
    o, xs, ms = small_slices()
    hb = batch.HeteroBatch([torch.from_numpy(oo) for oo in o], device=dev, graph=False)
    xt = [torch.from_numpy(x).to(dev) for x in xs]
    mt = [torch.from_numpy(m).to(dev) for m in ms]
    print("[case] batch schedules", flush=True)
    hb.run(xt, mt)
    ref = [g.clone() for g in hb.grad]
    for bpc, order in ((1, 0), (3, 1), (4, 0)):
        hb.set_schedule(bpc, order)
        hb.run(xt, mt)
        for s in range(len(o)):
            assert torch.equal(hb.grad[s], ref[s]), (bpc, order, s)
    print("[case] batch graph replays", flush=True)
    hb.use_graph = True
    for rep in range(4):
        # eager work between replays on the same stream
        for g in hb.grad:
            g.zero_()
        junk = torch.rand(1 << 20, device=dev).sum()
        hb.run(xt, mt)
        torch.cuda.synchronize()
        for s in range(len(o)):
            assert torch.equal(hb.grad[s], ref[s]), (rep, s)
        del junk


def test_batch_config3_against_plain(dev):
    """The 64-slice heterogeneous batch (sizes {256,384,512}^2, views U[96,1152], bins {512,736,1024},
    scan types {0,1,2}) equals 64 plain per-slice calls; 4 slices are also checked against the oracle."""
    opts, imgs = phantom.hetero_slices()
    xs = [torch.from_numpy(im)[None, None].to(dev) for im in imgs]
    ms = []
    for oo, x in zip(opts, xs):
        ms.append(ctlib_amd.projection(x * 1.1, torch.from_numpy(oo)))
    print("[case] config3 batch: build + tune", flush=True)
    hb = batch.HeteroBatch([torch.from_numpy(oo) for oo in opts], device=dev, graph=True)
    ys, grads = hb.run(xs, ms)
    hb.check()
    print("[case] config3 batch: compare", flush=True)
    for s, oo in enumerate(opts):
        y_ref, g_ref = ctlib_amd.residual_step(xs[s], ms[s], torch.from_numpy(oo))
        assert rel(ys[s].cpu().numpy(), y_ref.cpu().numpy()) < 1e-6, s
        assert rel(grads[s].cpu().numpy(), g_ref.cpu().numpy()) < 1e-6, s
    for s in (0, 17, 40, 63):
        o64 = opts[s].astype(np.float64)
        x64 = xs[s].cpu().numpy()[:, 0].astype(np.float64)
        y_or = O.apply("projection", o64, x64)
        g_or = O.apply("projection_t", o64, y_or - ms[s].cpu().numpy()[:, 0].astype(np.float64))
        assert rel(ys[s].cpu().numpy()[:, 0], y_or) < 1e-5, s
        assert rel(grads[s].cpu().numpy()[:, 0], g_or) < 1e-5, s


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_with_large_slices(dev, dtype):
    """A batch holding slices beyond 512 x 512 / 1279 bins (config 4's 1024^2, 1536-bin fan-ED geometry
    with fewer views, a 768 x 640 fan-EA slice with 1300 bins, a 1000 x 1000 parallel slice) together
    with small ones runs the large-line kernel variant for every slice and equals the plain per-slice
    calls; the small slices also match the oracle."""
    o = [
        phantom.fan_options(180, 1536, 1024, 1024, 0.0033, 0.0036, 5.95, 4.906, 0),
        phantom.fan_options(90, 1300, 768, 640, 0.004, 0.0006, 5.95, 4.906, 1, binshift=0.0002),
        phantom.para_options(60, 1100, 1000, 1000, 0.002, 0.002, binshift=0.001),
        phantom.fan_options(60, 47, 32, 32, 0.05, 0.06, 3.0, 2.0, 0),
        phantom.para_options(36, 53, 40, 40, 0.03, 0.03),
    ]
    rng = np.random.default_rng(78)
    xs, ms = [], []
    for oo in o:
        g = O.parse(oo)
        xs.append(torch.from_numpy(rng.random((1, 1, g["height"], g["width"])).astype(dtype)).to(dev))
        ms.append(torch.from_numpy(rng.random((1, 1, g["views"], g["dets"])).astype(dtype)).to(dev))
    print(f"[case] batch large slices {dtype.__name__}: build", flush=True)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    hb = batch.HeteroBatch([torch.from_numpy(oo) for oo in o], device=dev, dtype=tdt, graph=False)
    print(f"[case] batch large slices: run ({hb.info()['items']} items)", flush=True)
    ys, grads = hb.run(xs, ms)
    hb.check()
    # same arithmetic, different summation order (the plain calls split lines into chunks): fp64
    # rounding over 1024-pixel lines reaches ~1e-12
    tol_plain = 1e-6 if dtype == np.float32 else 1e-10
    for s, oo in enumerate(o):
        y_ref, g_ref = ctlib_amd.residual_step(xs[s], ms[s], torch.from_numpy(oo))
        assert rel(ys[s].cpu().numpy(), y_ref.cpu().numpy()) < tol_plain, s
        assert rel(grads[s].cpu().numpy(), g_ref.cpu().numpy()) < tol_plain, s
    for s in (3, 4):
        o64 = o[s].astype(np.float64)
        x64 = xs[s].cpu().numpy()[:, 0].astype(np.float64)
        y_or = O.apply("projection", o64, x64)
        g_or = O.apply("projection_t", o64, y_or - ms[s].cpu().numpy()[:, 0].astype(np.float64))
        tol = 1e-5 if dtype == np.float32 else 1e-7
        assert rel(ys[s].cpu().numpy()[:, 0], y_or) < tol, s
        assert rel(grads[s].cpu().numpy()[:, 0], g_or) < tol, s


def test_tuner_roundtrip(dev, tmp_path, monkeypatch):
    monkeypatch.setenv("CTLIB_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    o = phantom.fan_options(96, 128, 64, 64, 0.05, 0.06, 5.95, 4.906, 0)
    print("[case] tuner plan", flush=True)
    best, times = tune.tune_plan(o, batch=2, reps=2)
    assert best in (4, 8) and set(times) == {4, 8}
    assert tune.apply_cached(o) == best
    x = torch.rand(2, 1, 64, 64, device=dev)
    y = ctlib_amd.projection(x, torch.from_numpy(o))
    ref = O.apply("projection", o.astype(np.float64), x.cpu().numpy()[:, 0].astype(np.float64))
    assert rel(y.cpu().numpy()[:, 0], ref) < 1e-5
    print("[case] tuner batch", flush=True)
    so, xs, ms = small_slices()
    hb = batch.HeteroBatch([torch.from_numpy(oo) for oo in so], device=dev, graph=False)
    sched, t = tune.tune_batch(hb, reps=2)
    assert tune.cached_batch_schedule(hb.signature()) == sched
    hb2 = batch.HeteroBatch([torch.from_numpy(oo) for oo in so], device=dev, graph=False)
    assert (hb2.info()["blocks_per_cu"], hb2.info()["order"]) == sched


def test_batch_f64_matches_plain_and_oracle(dev):
    """The persistent batch kernel with fp64 data (every scan type, both sweep axes) equals the
    per-slice fused residual step to 1e-12 and the FP64 oracle to 1e-7."""
    o, xs, ms = small_slices()
    print("[case] batch small fp64", flush=True)
    hb = batch.HeteroBatch([torch.from_numpy(oo) for oo in o], device=dev, dtype=torch.float64, graph=False)
    xt = [torch.from_numpy(x.astype(np.float64)).to(dev) for x in xs]
    mt = [torch.from_numpy(m.astype(np.float64)).to(dev) for m in ms]
    ys, grads = hb.run(xt, mt)
    hb.check()
    for s, oo in enumerate(o):
        assert ys[s].dtype == torch.float64 and grads[s].dtype == torch.float64
        y_ref, g_ref = ctlib_amd.residual_step(xt[s], mt[s], torch.from_numpy(oo))
        assert rel(ys[s].cpu().numpy(), y_ref.cpu().numpy()) < 1e-12, s
        assert rel(grads[s].cpu().numpy(), g_ref.cpu().numpy()) < 1e-12, s
        o64 = oo.astype(np.float64)
        y_or = O.apply("projection", o64, xs[s][:, 0].astype(np.float64))
        g_or = O.apply("projection_t", o64, y_or - ms[s][:, 0].astype(np.float64))
        assert rel(ys[s].cpu().numpy()[:, 0], y_or) < 1e-7, s
        assert rel(grads[s].cpu().numpy()[:, 0], g_or) < 1e-7, s
