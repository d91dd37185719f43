"""The measured configurations at their stated batch sizes against the FP64 oracle
(oracle/ct_oracle.cpp), slices sampled where the oracle would take minutes:

  * fan-ED 512^2, 736 x 736, B=32 (the headline workload): slices 0 and 31 of projection
    and of the least-squares gradient A^T (A x - m) from the fused residual call;
  * fan-ED 512^2, 736 x 736, B=8: the last slice of every operator;
  * fan-EA 512^2, 1024 x 888, B=16: the last slice of projection, projection_t and
    backprojection_t;
  * fan-ED 1024^2, 1440 x 1536, B=4, Poisson noise I0=5e4: a checkpointed unrolled
    gradient descent (learnable step sizes, trained through forward + backward) against the
    oracle's own iterations on slice 0, and its peak activation memory.

Bar: fp32 data, relative L2 <= 1e-5 (max-abs printed)."""
import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import _C, autograd, phantom, recon
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def check(name, got, ref):
    err = rel_l2(got, ref)
    print(f"  {name}: rel_l2 {err:.3e}, max-abs {np.abs(got - ref).max():.3e}", flush=True)
    assert err <= F32_TOL, (name, err)


def test_metric_workload_b32_slices(dev):
    o = phantom.CONFIGS[1]["options"]
    ot = torch.from_numpy(o)
    o64 = o.astype(np.float64)
    B = 32
    x_true = phantom.phantom_batch(B, 512, 512, config=1, scale=0.02)
    xt = torch.from_numpy(x_true).to(dev)
    print("[case] fan-ED 512^2 B=32: projection and least-squares gradient, slices 0 and 31", flush=True)
    m = ctlib_amd.projection(xt, ot)
    x0 = (0.9 * xt).contiguous().requires_grad_(True)
    loss = autograd.least_squares(x0, m, ot)
    loss.backward()
    y = m.cpu().numpy()
    g = x0.grad.cpu().numpy()
    for s in (0, B - 1):
        y_ref = O.apply("projection", o64, x_true[s:s + 1, 0].astype(np.float64))
        check(f"projection slice {s}", y[s:s + 1, 0], y_ref)
        # gradient at x0 (fp32, as the GPU holds it) with the GPU's fp32 measurement m
        xs = x0.detach()[s:s + 1, 0].cpu().double().numpy()
        r = O.apply("projection", o64, xs) - y[s:s + 1, 0].astype(np.float64)
        check(f"least-squares gradient slice {s}", g[s:s + 1, 0], O.apply("projection_t", o64, r))


OPS = ("projection", "projection_t", "backprojection", "backprojection_t", "fbp")
IMG_IN = ("projection", "backprojection_t")


@pytest.mark.parametrize("cfg,ops", [(1, OPS), (2, ("projection", "projection_t", "backprojection_t"))])
def test_stated_batch_last_slice(dev, cfg, ops):
    o = phantom.CONFIGS[cfg]["options"]
    B = phantom.CONFIGS[cfg]["batch"]
    ot = torch.from_numpy(o)
    o64 = o.astype(np.float64)
    g = O.parse(o64)
    x = phantom.phantom_batch(B, g["height"], g["width"], config=cfg, scale=0.02)
    xt = torch.from_numpy(x).to(dev)
    yt = ctlib_amd.projection(xt, ot)
    y = yt.cpu().numpy()
    for name in ops:
        print(f"[case] config {cfg} B={B} {name}, last slice", flush=True)
        inp, arr = (xt, x) if name in IMG_IN else (yt, y)
        got = getattr(ctlib_amd, name)(inp, ot)[B - 1:B, 0].cpu().numpy()
        ref = O.apply(name, o64, arr[B - 1:B, 0].astype(np.float64))
        check(f"config {cfg} {name}", got, ref)


def test_config4_unrolled_reconstruction(dev):
    """Fan-ED 1024^2, 1440 views x 1536 bins, B=4, Poisson I0=5e4: K=5 checkpointed GD
    iterations with learnable step sizes.  The reconstruction of slice 0 equals the FP64
    oracle's 5 iterations from the same measurement, and a training step (forward + backward
    through all iterations) keeps only image-sized activations per iteration."""
    cfg = phantom.CONFIGS[4]
    o = cfg["options"]
    ot = torch.from_numpy(o)
    o64 = o.astype(np.float64)
    B, K = cfg["batch"], 5
    x_true = phantom.phantom_batch(B, 1024, 1024, config=4, scale=0.02)
    xt_true = torch.from_numpy(x_true).to(dev)
    clean = ctlib_amd.projection(xt_true, ot).double().cpu().numpy()
    y = np.stack([phantom.poisson_sinogram(clean[b, 0], 5e4, seed=1_000_000 + 4 + b) for b in range(B)])[:, None]
    y = y.astype(np.float32)
    yt = torch.from_numpy(y).to(dev)
    step = recon.lipschitz_step(ot, (1, 1, 1024, 1024), dev)
    model = recon.UnrolledGD(ot, iters=K, alpha0=step, checkpoint=True).to(dev)
    print("[case] config 4 unrolled GD, K=5, training step", flush=True)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    loss = ((model(yt) - xt_true) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    assert model.alpha.grad is not None and bool(torch.isfinite(model.alpha.grad).all())
    img = 4 * B * 1024 * 1024
    sino = 4 * B * 1440 * 1536
    ws = max(_C.workspace_bytes(ot, op, B, False, dev.index or 0) for op in (0, 1))
    # kept per iteration: the iterate (checkpoint input); transient: one iteration's
    # recomputation and its backward (two sinograms, two images) and one call's workspace
    bound = K * img + 2 * sino + 3 * img + ws
    print(f"  peak activation {peak / 1e6:.1f} MB, bound {bound / 1e6:.1f} MB "
          f"(an iteration's sinogram kept per step would add {K * sino / 1e6:.0f} MB)", flush=True)
    assert peak <= bound, (peak, bound)
    with torch.no_grad():
        xr = model(yt)[0:1, 0].double().cpu().numpy()
    print("[case] config 4 oracle iterations, slice 0", flush=True)
    xo = np.zeros((1, 1024, 1024))
    y0 = y[0:1, 0].astype(np.float64)
    a = model.alpha.detach().to(torch.float32).double().cpu().numpy()  # the step the GPU applies (fp32)
    for k in range(K):
        xo = xo - a[k] * O.apply("projection_t", o64, O.apply("projection", o64, xo) - y0)
    check("config 4 reconstruction slice 0", xr, xo)
