"""The projector family as torch.library custom operators (namespace `ctlib_amd`).

    torch.ops.ctlib_amd.projection(image, options)          options: list of floats
    torch.ops.ctlib_amd.projection_t(sino, options)
    torch.ops.ctlib_amd.backprojection(sino, options)
    torch.ops.ctlib_amd.backprojection_t(image, options)
    torch.ops.ctlib_amd.backprojection_t_true(image, options)
    torch.ops.ctlib_amd.fbp(sino, options)
    torch.ops.ctlib_amd.fbp_t(image, options)
    torch.ops.ctlib_amd.backprojection_sv(sino, options)
    torch.ops.ctlib_amd.residual_step(image, measured, options) -> (A x, A^T (A x - m))

Each operator has a fake (meta) implementation that derives the output shape
from the options, so torch.compile traces them with fullgraph=True and
FakeTensor / AOT dispatch see exact shapes; the differentiable ones register
their matched adjoint as the backward (the same pairs as ctlib_amd.autograd;
for scan_type 1 the reference's backprojection_t carries dImg / c_k per bin,
fan_ea_kernel.cu:590-607, so its backward applies that factor before the
backprojection).  The options travel as a list of Python floats (a constant
of the traced graph), not as a tensor whose values a fake tensor cannot see.
The reference registers nothing of the kind (ctlib.cpp:106-114 is a plain
pybind module).
"""
import functools

import numpy as np
import torch

from . import _C, abi


@functools.lru_cache(maxsize=256)
def _geom(opts):
    """Geometry of an options tuple through the C ABI on host memory (usable under fake-tensor tracing)."""
    g = abi.parse_options(np.asarray(opts, dtype=np.float64))
    return {"views": g.views, "dets": g.dets, "width": g.width, "height": g.height, "scan_type": g.scan_type}


def _t(opts):
    return torch.tensor(opts, dtype=torch.float64)


def _img_shape(x, opts):
    g = _geom(tuple(opts))
    return (x.shape[0], 1, g["height"], g["width"])


def _sino_shape(x, opts):
    g = _geom(tuple(opts))
    return (x.shape[0], 1, g["views"], g["dets"])


def as_list(options):
    """The options vector as a list of floats (host tensors and arrays accepted)."""
    if isinstance(options, torch.Tensor):
        return [float(v) for v in options.detach().to("cpu", torch.float64).tolist()]
    return [float(v) for v in np.asarray(options, dtype=np.float64).ravel()]


def _define(name, run, out_shape):
    @torch.library.custom_op(f"ctlib_amd::{name}", mutates_args=())
    def op(x: torch.Tensor, options: list[float]) -> torch.Tensor:
        return run(x.contiguous(), _t(options))

    @op.register_fake
    def _(x, options):
        return x.new_empty(out_shape(x, options))

    return op


projection = _define("projection", _C.projection, _sino_shape)
projection_t = _define("projection_t", _C.projection_t, _img_shape)
backprojection = _define("backprojection", _C.backprojection, _img_shape)
backprojection_t = _define("backprojection_t", _C.backprojection_t, _sino_shape)
backprojection_t_true = _define("backprojection_t_true", _C.backprojection_t_true, _sino_shape)
fbp = _define("fbp", _C.fbp, _img_shape)
fbp_t = _define("fbp_t", _C.fbp_t, _sino_shape)


@torch.library.custom_op("ctlib_amd::backprojection_sv", mutates_args=())
def backprojection_sv(x: torch.Tensor, options: list[float]) -> torch.Tensor:
    return _C.backprojection_sv(x.contiguous(), _t(options))


@backprojection_sv.register_fake
def _(x, options):
    g = _geom(tuple(options))
    if g["scan_type"] != 0:
        raise ValueError("backprojection_sv supports scan_type 0 (fan beam, equal distance) only")
    return x.new_empty((x.shape[0], g["views"], g["height"], g["width"]))


@torch.library.custom_op("ctlib_amd::residual_step", mutates_args=())
def residual_step(x: torch.Tensor, m: torch.Tensor, options: list[float]) -> tuple[torch.Tensor, torch.Tensor]:
    y, grad = _C.residual_step(x.contiguous(), m.contiguous(), _t(options))
    return y, grad


@residual_step.register_fake
def _(x, m, options):
    return x.new_empty(_sino_shape(x, options)), x.new_empty(x.shape)


# ---- autograd: each operator's backward is its matched adjoint --------------

def _setup(ctx, inputs, output):
    ctx.options = inputs[1]


def _register_adjoint(op, adjoint):
    def backward(ctx, grad):
        return adjoint(grad.contiguous(), ctx.options), None

    op.register_autograd(backward, setup_context=_setup)


_register_adjoint(projection, lambda g, o: torch.ops.ctlib_amd.projection_t(g, o))
_register_adjoint(projection_t, lambda g, o: torch.ops.ctlib_amd.projection(g, o))
_register_adjoint(backprojection, lambda g, o: torch.ops.ctlib_amd.backprojection_t_true(g, o))
_register_adjoint(backprojection_t_true, lambda g, o: torch.ops.ctlib_amd.backprojection(g, o))
_register_adjoint(fbp, lambda g, o: torch.ops.ctlib_amd.fbp_t(g, o))
_register_adjoint(fbp_t, lambda g, o: torch.ops.ctlib_amd.fbp(g, o))


@functools.lru_cache(maxsize=64)
def _bt_factor_host(opts):
    """Per (view, output bin) factor of the reference backprojection_t over the exact adjoint (fan-EA)."""
    o = np.asarray(opts, dtype=np.float64)
    g = abi.parse_options(o)
    post = abi.plan_export(o, "post_bprj_t").reshape(g.views, g.dets)
    rev = abi.plan_export(o, "views")["rev"].astype(bool)
    return np.ascontiguousarray(np.where(rev[:, None], post[:, ::-1], post))


def _bt_backward(ctx, grad):
    g = grad.contiguous()
    if int(ctx.options[-1]) == 1:
        f = torch.as_tensor(_bt_factor_host(tuple(ctx.options)), dtype=g.dtype, device=g.device)
        g = (g * f.view(1, 1, *f.shape)).contiguous()
    return torch.ops.ctlib_amd.backprojection(g, ctx.options), None


backprojection_t.register_autograd(_bt_backward, setup_context=_setup)
