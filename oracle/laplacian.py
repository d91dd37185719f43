"""FP64 numpy restatement of the reference's kNN graph weights (`laplacian`).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg,
never by the ctlib_amd package.

laplacian_cuda_kernel.cu:46-128, for points x [N, d] and k:
  norm2[i]  = sum_j x[i, j]^2                                   (:9-22)
  dist[i,j] = norm2[i] + norm2[j] - 2 x_i . x_j                 (:24-43 add_norm, :85-107 GEMM with beta 1)
  (coef, cols) = topk(dist, k, dim=1, largest=False)            (:112) ascending; ties by column here
  median    = lower median of all N*k values of coef            (:115, torch.median)
  coef      = exp(-coef / median), then divided by its row sum  (:116-119)
  W         = sparse COO (N, N), row i -> cols[i], float32 values (:120-126)
Returned as (cols int64 [N, k], vals float32 [N, k]); row indices are implicit.
"""
import numpy as np


def distances(x):
    x = np.asarray(x, dtype=np.float64)
    n2 = (x * x).sum(1)
    return n2[:, None] + n2[None, :] - 2.0 * (x @ x.T)


def laplacian(x, k, dist=None):
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    if not 1 <= k <= N:
        raise ValueError("k must be in [1, N]")
    d = distances(x) if dist is None else dist
    cols = np.empty((N, k), dtype=np.int64)
    coef = np.empty((N, k), dtype=np.float64)
    idx = np.arange(N)
    for i in range(N):
        order = np.lexsort((idx, d[i]))[:k]
        cols[i] = order
        coef[i] = d[i, order]
    flat = np.sort(coef.ravel())
    median = flat[(flat.size - 1) // 2]
    w = np.exp(-coef / median)
    w = w / w.sum(1, keepdims=True)
    return cols, w.astype(np.float32)
