"""Differentiation on the GPU: autograd wrappers whose backward is the matched
adjoint kernel, checked by central finite differences in FP64, by the adjoint
dot-product identity and against the FP64 oracle's transpose; checkpointed
unrolled reconstruction."""
import math

import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import autograd, phantom, recon
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small(st, V=40, D=29, W=16, H=14):
    if st == 2:
        return phantom.para_options(V, D, W, H, 0.1, 0.1, Ang0=0.2, dAng=math.pi / V, binshift=0.01)
    return phantom.fan_options(V, D, W, H, 0.1, 0.03 if st == 1 else 0.12, 3.0, 2.0, st, Ang0=0.2, binshift=0.01)


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("name", ["projection", "projection_t", "backprojection", "backprojection_t", "fbp"])
def test_gradcheck_f64(dev, st, name):
    """torch.autograd.gradcheck (FD in fp64) of every wrapper."""
    o = torch.from_numpy(small(st).astype(np.float64))
    g = O.parse(o.numpy())
    img_in = name in ("projection", "backprojection_t")
    shape = (1, 1, g["height"], g["width"]) if img_in else (1, 1, g["views"], g["dets"])
    x = torch.rand(shape, dtype=torch.float64, device=dev, generator=None, requires_grad=True)
    fn = getattr(autograd, name)
    print(f"[case] gradcheck st={st} {name}", flush=True)
    # linear operators: the FD Jacobian is exact up to rounding; sample a few output directions
    assert torch.autograd.gradcheck(lambda t: fn(t, o), (x,), eps=1e-3, atol=1e-7, rtol=1e-6, fast_mode=True)


@pytest.mark.parametrize("st", [0, 1, 2])
def test_backward_is_oracle_transpose(dev, st):
    """The backward of each wrapper equals the FP64 oracle's transpose applied to the cotangent."""
    o = small(st).astype(np.float64)
    g = O.parse(o)
    rng = np.random.default_rng(20 + st)
    x = torch.from_numpy(rng.random((2, 1, g["height"], g["width"]))).to(dev).requires_grad_(True)
    cot = rng.random((2, 1, g["views"], g["dets"]))
    print(f"[case] backward vs oracle st={st}", flush=True)
    y = autograd.projection(x, torch.from_numpy(o))
    y.backward(torch.from_numpy(cot).to(dev))
    ref = O.apply("projection_t", o, cot[:, 0])[:, None]
    got = x.grad.cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-7
    # backprojection: backward = exact adjoint (also for EA, where the reference's backprojection_t is not)
    s = torch.from_numpy(cot).to(dev).requires_grad_(True)
    img_cot = rng.random((2, 1, g["height"], g["width"]))
    autograd.backprojection(s, torch.from_numpy(o)).backward(torch.from_numpy(img_cot).to(dev))
    ref = O.apply("backprojection_t_true", o, img_cot[:, 0])[:, None]
    assert np.linalg.norm(s.grad.cpu().numpy() - ref) / np.linalg.norm(ref) < 1e-7


def test_least_squares_fd_config2(dev):
    """Fan-EA 512^2, 1024 views x 888 bins: d/dx ||Ax - y||^2 by central differences (fp64, h = 1e-3)
    on 64 seeded pixels against the hand-written adjoint, with y a Poisson-noised sinogram (I0 = 1e5)."""
    o = phantom.CONFIGS[2]["options"].astype(np.float64)
    ot = torch.from_numpy(o)
    x_true = phantom.phantom_batch(1, 512, 512, config=2, scale=0.02).astype(np.float64)
    clean = O.apply("projection", o, x_true[:, 0])
    y = phantom.poisson_sinogram(clean, 1e5, seed=1_000_000 + 2).astype(np.float64)[:, None]
    xt = torch.from_numpy(0.5 * x_true).to(dev).requires_grad_(True)
    yt = torch.from_numpy(y).to(dev)
    print("[case] FD config2", flush=True)
    loss = ((autograd.projection(xt, ot) - yt) ** 2).sum()
    loss.backward()
    grad = xt.grad.detach()
    rng = np.random.default_rng(2024)
    pix = rng.integers(0, 512, size=(64, 2))
    h = 1e-3
    with torch.no_grad():
        for r, c in pix:
            xp = xt.detach().clone()
            xp[0, 0, r, c] += h
            xm = xt.detach().clone()
            xm[0, 0, r, c] -= h
            lp = ((ctlib_amd.projection(xp, ot) - yt) ** 2).sum().item()
            lm = ((ctlib_amd.projection(xm, ot) - yt) ** 2).sum().item()
            fd = (lp - lm) / (2 * h)
            an = grad[0, 0, r, c].item()
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-2 * float(grad.abs().max())), (r, c, fd, an)


def test_adjoint_dot_config2(dev):
    o = torch.from_numpy(phantom.CONFIGS[2]["options"])
    x = torch.rand(2, 1, 512, 512, device=dev)
    y = torch.rand(2, 1, 1024, 888, device=dev)
    print("[case] adjoint dot config2", flush=True)
    lhs = float((ctlib_amd.projection(x, o).double() * y.double()).sum())
    rhs = float((x.double() * ctlib_amd.projection_t(y, o).double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_unrolled_gd_matches_oracle_and_checkpoints(dev):
    """10 unrolled GD steps (fan-ED 64^2) equal the FP64 oracle's iteration; dL/dalpha by FD;
    checkpointing keeps less activation memory."""
    o = phantom.fan_options(90, 96, 64, 64, 0.0264, 0.0288, 5.95, 4.906, 0).astype(np.float64)
    ot = torch.from_numpy(o)
    x_true = phantom.phantom_batch(2, 64, 64, config=4, scale=0.02).astype(np.float64)
    y = O.apply("projection", o, x_true[:, 0])
    step = recon.lipschitz_step(ot, (2, 1, 64, 64), dev, dtype=torch.float64)
    K = 10
    model = recon.UnrolledGD(ot, iters=K, alpha0=step, checkpoint=True).to(dev)
    yt = torch.from_numpy(y[:, None]).to(dev)
    print("[case] unrolled GD", flush=True)
    xr = model(yt)
    # FP64 oracle run of the same iteration
    xo = np.zeros_like(x_true[:, 0])
    for _ in range(K):
        xo = xo - step * O.apply("projection_t", o, O.apply("projection", o, xo) - y)
    err = np.linalg.norm(xr.detach().cpu().numpy()[:, 0] - xo) / np.linalg.norm(xo)
    assert err < 1e-10, err
    # gradient w.r.t. the step sizes against central differences
    loss = ((xr - torch.from_numpy(x_true).to(dev)) ** 2).sum()
    loss.backward()
    k = 3
    h = 1e-6 * step
    with torch.no_grad():
        for sgn, store in ((1, "p"), (-1, "m")):
            model.alpha[k] += sgn * h
            val = ((model(yt) - torch.from_numpy(x_true).to(dev)) ** 2).sum().item()
            model.alpha[k] -= sgn * h
            if store == "p":
                lp = val
            else:
                lm = val
    fd = (lp - lm) / (2 * h)
    assert abs(fd - model.alpha.grad[k].item()) <= 1e-5 * abs(fd), (fd, model.alpha.grad[k].item())


def test_checkpointing_memory(dev):
    """With a per-iteration regularizer network, checkpointing keeps only the iterates:
    peak memory drops and the gradients are unchanged."""
    o = torch.from_numpy(phantom.fan_options(180, 192, 128, 128, 0.0132, 0.0144, 5.95, 4.906, 0))
    y = torch.rand(4, 1, 180, 192, device=dev)
    torch.manual_seed(0)
    reg = torch.nn.Sequential(torch.nn.Conv2d(1, 32, 3, padding=1), torch.nn.ReLU(),
                              torch.nn.Conv2d(32, 32, 3, padding=1), torch.nn.ReLU(),
                              torch.nn.Conv2d(32, 1, 3, padding=1)).to(dev)
    peaks, grads = {}, {}
    print("[case] checkpoint memory", flush=True)
    for ck in (False, True):
        model = recon.UnrolledGD(o, iters=12, alpha0=1e-3, checkpoint=ck, regularizer=reg).to(dev)
        reg.zero_grad()
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        model(y).square().sum().backward()
        torch.cuda.synchronize()
        peaks[ck] = torch.cuda.max_memory_allocated() - base
        grads[ck] = (model.alpha.grad.clone(), reg[0].weight.grad.clone())
    assert peaks[True] < 0.6 * peaks[False], peaks
    torch.testing.assert_close(grads[True][0], grads[False][0], rtol=1e-4, atol=0)
    torch.testing.assert_close(grads[True][1], grads[False][1], rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("chunks", [0, 1, 3])
@pytest.mark.parametrize("st", [0, 1, 2])
def test_least_squares_loss_and_grad(dev, st, chunks, dtype):
    """least_squares: the loss 0.5 * ||A x - m||^2 (from the library's per-row FP64 sums) and its
    gradient A^T (A x - m) against the FP64 oracle, for the forward fused with the residual tables
    (1 line chunk) and the split forward (automatic and 3 chunks)."""
    from ctlib_amd import _C
    o = small(st)
    g = O.parse(o)
    rng = np.random.default_rng(70 + st)
    x = rng.random((2, 1, g["height"], g["width"]))
    m = rng.random((2, 1, g["views"], g["dets"]))
    ot = torch.from_numpy(o)
    print(f"[case] least squares st={st} chunks={chunks} {dtype}", flush=True)
    xt = torch.from_numpy(x).to(dev, dtype).requires_grad_(True)
    mt = torch.from_numpy(m).to(dev, dtype)
    try:
        _C.set_plan_chunks(ot, dev.index or 0, chunks, 0)
        loss = autograd.least_squares(xt, mt, ot)
        loss.backward()
        y, _ = ctlib_amd.residual_step(xt.detach(), mt, ot)
    finally:
        _C.set_plan_chunks(ot, dev.index or 0, 0, 0)
    xr = xt.detach().cpu().double().numpy()[:, 0]
    mr = mt.cpu().double().numpy()[:, 0]
    r = O.apply("projection", o.astype(np.float64), xr) - mr
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    assert loss.dtype == dtype
    assert abs(loss.item() - 0.5 * float((r * r).sum())) <= tol * 0.5 * float((r * r).sum())
    # the same sum from the stored A x (rounded to the data type) agrees to that rounding
    r_dev = (y.double() - mt.double()).cpu().numpy()
    assert abs(loss.item() - 0.5 * float((r_dev * r_dev).sum())) <= tol * 0.5 * float((r * r).sum())
    g_ref = O.apply("projection_t", o.astype(np.float64), r)
    got = xt.grad.cpu().double().numpy()[:, 0]
    assert np.linalg.norm(got - g_ref) / np.linalg.norm(g_ref) <= (1e-5 if dtype == torch.float32 else 1e-7)
