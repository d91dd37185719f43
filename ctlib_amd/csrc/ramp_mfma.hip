// RL ramp filter of fbp as a dense Toeplitz GEMM on FP64 MFMA (gfx950).
//
// The reference filters every pre-weighted sinogram row with
// torch::conv2d(row, h, padding = D-1) (fan_ed_kernel.cu:1368,
// fan_ea_kernel.cu:1093, para_kernel.cu:836), a cross-correlation
//   q[j] = sum_m row[m] * h(m - j),  h(n) = RL taps (fan_ed_kernel.cu:925-943).
// Over all B*V rows this is [B*V x D] x [D x D] with the Toeplitz matrix
// T[m][j] = h(m - j), run here as v_mfma_f64_16x16x4f64 tiles.  The taps are
// staged once per block in LDS; the pre-weight is fused into the A-operand
// load and the post-scale into the store.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctlib {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Block: 4 waves, 64 rows x 64 columns; each wave owns 16 rows x 64 columns
// (four 16x16 accumulators).  Operand layout of v_mfma_f64_16x16x4f64 (checked
// on gfx950): A[i][k] and B[k][j] are held by lane i + 16k (resp. j + 16k);
// accumulator element r of lane l is D[l/16 + 4r][l%16].
template <typename TIn, typename TOut>
__global__ __launch_bounds__(256) void k_ramp(const TIn* __restrict__ in, TOut* __restrict__ out, long rows, int D,
                                             const double* __restrict__ taps, const double* __restrict__ pre,
                                             const double* __restrict__ post, double scale) {
    extern __shared__ double sh_taps[];  // [2D-1]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 2 * D - 1; i += 256) sh_taps[i] = taps[i];
    __syncthreads();
    const long r0 = (long)blockIdx.y * 64 + wave * 16;
    const int j0 = blockIdx.x * 64;
    const int li = lane & 15, kk = lane >> 4;
    const long ra = r0 + li;
    const bool row_ok = ra < rows;
    const TIn* arow = in + (row_ok ? ra : 0) * (long)D;
    dbl4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < D; k0 += 4) {
        const int k = k0 + kk;
        double a = 0.0;
        if (row_ok && k < D) {
            a = (double)arow[k];
            if (pre) a *= pre[k];
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int j = j0 + 16 * t + li;
            const double bv = (k < D && j < D) ? sh_taps[k - j + D - 1] : 0.0;
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = j0 + 16 * t + li;
        if (col >= D) continue;
        const double f = post ? post[col] * scale : scale;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long row = r0 + kk + 4 * r;
            if (row < rows) out[row * D + col] = (TOut)(acc[t][r] * f);
        }
    }
}

template <typename TIn, typename TOut>
hipError_t ramp_t(const DevPlan& P, const void* in, void* out, long rows, const double* pre, const double* post,
                  double scale, hipStream_t s) {
    const int D = P.D;
    dim3 grid((D + 63) / 64, (unsigned)((rows + 63) / 64));
    const size_t lds = (size_t)(2 * D - 1) * sizeof(double);
    auto kern = k_ramp<TIn, TOut>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, (const TIn*)in, (TOut*)out, rows, D, P.taps, pre, post, scale);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ramp(const DevPlan& P, int in_dtype, const void* in, int out_dtype, void* out, long rows,
                       const double* pre, const double* post, double scale, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (in_dtype == DT_F32 && out_dtype == DT_F64) return ramp_t<float, double>(P, in, out, rows, pre, post, scale, s);
    if (in_dtype == DT_F64 && out_dtype == DT_F64) return ramp_t<double, double>(P, in, out, rows, pre, post, scale, s);
    if (in_dtype == DT_F32 && out_dtype == DT_F32) return ramp_t<float, float>(P, in, out, rows, pre, post, scale, s);
    return ramp_t<double, float>(P, in, out, rows, pre, post, scale, s);
}

}  // namespace ctlib
