"""Unrolled iterative reconstruction on the differentiable projectors.

UnrolledGD runs K steps of gradient descent on the least-squares data term,
    x_{k+1} = x_k - alpha_k * A^T (A x_k - y),
with one learnable step size alpha_k per iteration.  With checkpoint=True each
iteration is an activation checkpoint: only the iterates x_k are kept, and the
forward projections of an iteration are recomputed during backward, so memory
grows with K * |x| instead of K * (|x| + |A x|).
"""
import torch
from torch.utils.checkpoint import checkpoint as _checkpoint

from . import autograd


def _host_options(options):
    if isinstance(options, torch.Tensor):
        return options.detach().to("cpu")
    return torch.as_tensor(options)


def gd_step(x, y, alpha, options):
    r = autograd.projection(x, options) - y
    return x - alpha * autograd.projection_t(r, options)


def _iteration(x, y, alpha, options, reg):
    x = gd_step(x, y, alpha, options)
    if reg is not None:
        x = x + reg(x)
    return x


class UnrolledGD(torch.nn.Module):
    """K unrolled gradient steps; `regularizer` (optional nn.Module, image -> image)
    is added after each step, as in learned unrolled reconstruction networks.
    The pure data-term iteration only keeps image-sized tensors per step either
    way; checkpointing pays off once a regularizer network holds activations."""

    def __init__(self, options, iters=50, alpha0=1e-3, checkpoint=True, regularizer=None):
        super().__init__()
        self.options = _host_options(options)
        self.iters = iters
        self.checkpoint = checkpoint
        self.alpha = torch.nn.Parameter(torch.full((iters,), float(alpha0), dtype=torch.float64))
        self.regularizer = regularizer

    def forward(self, y, x0=None):
        g = self.options
        H, W = int(g[3]), int(g[2])
        x = x0 if x0 is not None else torch.zeros(y.shape[0], 1, H, W, device=y.device, dtype=y.dtype)
        for k in range(self.iters):
            a = self.alpha[k].to(y.dtype)
            if self.checkpoint and torch.is_grad_enabled():
                x = _checkpoint(_iteration, x, y, a, self.options, self.regularizer, use_reentrant=False)
            else:
                x = _iteration(x, y, a, self.options, self.regularizer)
        return x


def lipschitz_step(options, shape, device, dtype=torch.float32, iters=20, seed=0):
    """1 / ||A^T A|| by power iteration: a stable fixed step size for gradient descent."""
    gen = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.rand(shape, generator=gen, dtype=dtype).to(device)
    from . import _C
    opts = _host_options(options)
    lam = 1.0
    for _ in range(iters):
        x = x / torch.linalg.vector_norm(x)
        z = _C.projection_t(_C.projection(x, opts), opts)
        lam = float(torch.linalg.vector_norm(z))
        x = z
    return 1.0 / lam
