"""Differentiable wrappers: torch.autograd.Function per operator, with the
matched adjoint kernel as the hand-written backward.

    Projection       forward A          backward A^T   (projection_t)
    ProjectionT      forward A^T        backward A
    Backprojection   forward B          backward B^T   (exact adjoint, also for fan-EA)
    BackprojectionT  forward the reference's backprojection_t; backward its exact adjoint
    FBP              forward fbp        backward fbp_t
    LeastSquares     0.5 * ||A x - m||^2: the loss (per-row FP64 sums) and A^T (A x - m) from one call

The reference ships no autograd layer; callers wrapped projection /
projection_t themselves.  Note that for scan_type 1 the reference's
backprojection_t is not the adjoint of backprojection (it carries dImg / c_k,
fan_ea_kernel.cu:590-607), so Backprojection's backward uses the exact adjoint.
"""
import numpy as np
import torch

from . import _C, abi

_opt_cache = {}


def _host_options(options):
    if isinstance(options, torch.Tensor):
        return options.detach().to("cpu", torch.float64)
    return torch.as_tensor(np.asarray(options, dtype=np.float64))


class Projection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, options):
        ctx.options = options
        return _C.projection(image.contiguous(), options)

    @staticmethod
    def backward(ctx, grad):
        return _C.projection_t(grad.contiguous(), ctx.options), None


class ProjectionT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, options):
        ctx.options = options
        return _C.projection_t(sino.contiguous(), options)

    @staticmethod
    def backward(ctx, grad):
        return _C.projection(grad.contiguous(), ctx.options), None


class Backprojection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, options):
        ctx.options = options
        return _C.backprojection(sino.contiguous(), options)

    @staticmethod
    def backward(ctx, grad):
        return _C.backprojection_t_true(grad.contiguous(), ctx.options), None


def _bt_factor(options, device, dtype):
    """Per (view, output bin) factor of the reference backprojection_t over the exact adjoint."""
    host = _host_options(options)
    key = (tuple(host.tolist()), str(device), dtype)
    f = _opt_cache.get(key)
    if f is None:
        g = abi.parse_options(host.numpy())
        V, D = g.views, g.dets
        post = abi.plan_export(host.numpy(), "post_bprj_t").reshape(V, D)
        rev = abi.plan_export(host.numpy(), "views")["rev"].astype(bool)
        post = np.where(rev[:, None], post[:, ::-1], post)
        f = torch.as_tensor(np.ascontiguousarray(post), dtype=dtype, device=device).view(1, 1, V, D)
        _opt_cache[key] = f
    return f


class BackprojectionT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, options):
        ctx.options = options
        return _C.backprojection_t(image.contiguous(), options)

    @staticmethod
    def backward(ctx, grad):
        g = grad.contiguous()
        if int(_host_options(ctx.options)[-1]) == 1:
            g = (g * _bt_factor(ctx.options, g.device, g.dtype)).contiguous()
        return _C.backprojection(g, ctx.options), None


class FBP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, options):
        ctx.options = options
        return _C.fbp(sino.contiguous(), options)

    @staticmethod
    def backward(ctx, grad):
        return _C.fbp_t(grad.contiguous(), ctx.options), None


class LeastSquares(torch.autograd.Function):
    """0.5 * sum (A x - m)^2 over the batch; the gradient A^T (A x - m) comes from the same fused call."""

    @staticmethod
    def forward(ctx, image, measured, options):
        # per-row sums of (A x - m)^2 come from the same call (FP64), so A x is never stored
        rows, grad = _C.residual_loss_step(image.contiguous(), measured.contiguous(), options)
        ctx.save_for_backward(grad)
        return (0.5 * rows.sum()).to(image.dtype)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None


def projection(image, options):
    return Projection.apply(image, options)


def projection_t(sino, options):
    return ProjectionT.apply(sino, options)


def backprojection(sino, options):
    return Backprojection.apply(sino, options)


def backprojection_t(image, options):
    return BackprojectionT.apply(image, options)


def fbp(sino, options):
    return FBP.apply(sino, options)


def least_squares(image, measured, options):
    return LeastSquares.apply(image, measured, options)
