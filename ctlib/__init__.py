"""`import ctlib`: the reference module's Python surface, served by the MI355X library.

The reference builds one extension module named `ctlib` (setup.py:9) that exports seven
functions (ctlib.cpp:106-114).  This package exports the same seven names with the same
positional signatures, so code written against the reference runs unchanged:

    import ctlib
    sino = ctlib.projection(image, options)        # B x 1 x H x W -> B x 1 x V x D
    img  = ctlib.projection_t(sino, options)
    img  = ctlib.backprojection(sino, options)
    sino = ctlib.backprojection_t(image, options)
    imgs = ctlib.backprojection_sv(sino, options)  # B x V x H x W, scan_type 0
    img  = ctlib.fbp(sino, options)
    W    = ctlib.laplacian(points, k)              # sparse (N, N)

Each name is the function of `ctlib_amd` (hand-written gfx950 HIP kernels behind a C ABI);
the extras of that package (autograd wrappers, the fused residual step, heterogeneous
batches, the tuner) stay under `ctlib_amd`.
"""
from ctlib_amd import (backprojection, backprojection_sv, backprojection_t, fbp, laplacian, projection,  # noqa: F401
                       projection_t)

__version__ = "0.2.0"

__all__ = ["projection", "projection_t", "backprojection_t", "backprojection", "backprojection_sv", "fbp",
           "laplacian"]
