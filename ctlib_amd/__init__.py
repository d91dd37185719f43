"""ctlib_amd: MI355X-native distance-driven CT projectors with CTLIB's API.

Drop-in for the reference module `ctlib` (ctlib.cpp:106-114):

    import ctlib_amd as ctlib
    sino = ctlib.projection(image, options)          # B x 1 x H x W -> B x 1 x V x D
    img  = ctlib.projection_t(sino, options)         # exact adjoint of projection
    img  = ctlib.backprojection(sino, options)       # pixel-normalised, distance-weighted
    sino = ctlib.backprojection_t(image, options)    # transpose of backprojection (reference form)
    img  = ctlib.fbp(sino, options)                  # RL-filtered backprojection
    imgs = ctlib.backprojection_sv(sino, options)    # per-view backprojection, B x V x H x W (fan-ED)
    W    = ctlib.laplacian(points, k)                # kNN graph weights, sparse N x N

`options` is the reference's vector: 12 entries for fan beam
[views, dets, width, height, dImg, dDet, Ang0, dAng, s2r, d2r, binshift, scan_type]
and >= 9 for parallel beam [views, dets, width, height, dImg, dDet, Ang0, dAng,
binshift, ..., scan_type] (scan_type is always the last entry).  It may be a
host tensor (no device synchronisation) or a device tensor, float32 or float64.

All ops run hand-written gfx950 HIP kernels (libctlib_amd.so) on the current
HIP stream.  There is no CPU fallback: importing this package fails if the
extension is not built.
"""
import torch  # noqa: F401  (loads the HIP runtime the extension links against)

try:
    from . import _C
except ImportError as exc:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "ctlib_amd: the HIP extension (ctlib_amd/_C.so + libctlib_amd.so) is missing or failed to load; "
        "build it with `python -m ctlib_amd.build`") from exc

from . import ops  # noqa: E402,F401  (registers the torch.ops.ctlib_amd operators)

__version__ = "0.3.0"


def projection(image, options):
    """Distance-driven projection A (ctlib.cpp:56-69)."""
    return _C.projection(image, options)


def projection_t(projection, options):
    """Transpose of projection, A^T (ctlib.cpp:71-84)."""
    return _C.projection_t(projection, options)


def backprojection(projection, options):
    """Pixel-normalised, distance-weighted backprojection B (ctlib.cpp:30-43)."""
    return _C.backprojection(projection, options)


def backprojection_t(image, options):
    """Transpose of backprojection as the reference computes it (ctlib.cpp:15-28).

    For scan_type 1 (fan equal-angle) the reference multiplies by an extra
    dImg / c_k per bin (fan_ea_kernel.cu:590-607), so this op is not the exact
    adjoint of `backprojection` there; `backprojection_t_true` is.
    """
    return _C.backprojection_t(image, options)


def fbp(projection, options):
    """Filtered backprojection with the RL ramp filter (ctlib.cpp:86-99)."""
    return _C.fbp(projection, options)


def backprojection_t_true(image, options):
    """Exact adjoint of `backprojection` for every scan_type (no reference counterpart)."""
    return _C.backprojection_t_true(image, options)


def fbp_t(image, options):
    """Exact adjoint of `fbp` (no reference counterpart)."""
    return _C.fbp_t(image, options)


def residual_step(image, measured, options):
    """(A x, A^T (A x - m)) in one call: the data-term step of iterative reconstruction."""
    y, grad = _C.residual_step(image, measured, options)
    return y, grad


def parse_options(options):
    """Geometry the library reads from an options vector (reference indexing)."""
    if not isinstance(options, torch.Tensor):
        options = torch.as_tensor(options, dtype=torch.float64)
    return _C.parse_options(options)


def backprojection_sv(projection, options):
    """Per-view backprojection (ctlib.cpp:45-54, fan_ed_kernel.cu:945-1161): B x 1 x V x D -> B x V x H x W.

    Image v holds view v's term of `backprojection`, so the sum over dim 1 equals
    `backprojection(projection, options)[:, 0]`.  Like the reference, only
    scan_type 0 (fan beam, equal distance) is accepted; other scan types raise
    ValueError (the reference exits the process).  The output is B*V*H*W values:
    6.2 GB in float32 at B=8, 736 views, 512 x 512.
    """
    return _C.backprojection_sv(projection, options)


def laplacian(input, k):
    """kNN graph weights of the MAGIC network (ctlib.cpp:101-104, laplacian_cuda_kernel.cu:46-128).

    input: N x d points (float32 or float64, contiguous, on the GPU).  Returns the sparse COO
    (N, N) float32 tensor whose row i holds, at the k points nearest to point i by squared
    distance (itself included), exp(-dist / median) normalised to sum 1, median being the lower
    median of all N*k selected distances.  The N x N distance matrix is never formed (large d or
    k compute it in chunks of rows).  Supported: any d, k <= 4096.
    """
    return _C.laplacian(input, int(k))


__all__ = ["projection", "projection_t", "backprojection", "backprojection_t", "fbp", "backprojection_t_true",
           "fbp_t", "residual_step", "parse_options", "backprojection_sv", "laplacian"]
