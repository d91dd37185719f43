// RL ramp filter of fbp as a dense Toeplitz GEMM on FP64 MFMA (gfx950).
//
// The reference filters every pre-weighted sinogram row with
// torch::conv2d(row, h, padding = D-1) (fan_ed_kernel.cu:1368,
// fan_ea_kernel.cu:1093, para_kernel.cu:836), a cross-correlation
//   q[j] = sum_m row[m] * h(m - j),  h(n) = RL taps (fan_ed_kernel.cu:925-943).
// Over all B*V rows this is [B*V x D] x [D x D] with the Toeplitz matrix
// T[m][j] = h(m - j); since h vanishes at even offsets except 0 it runs as two
// half-size GEMMs (one per output-column parity) on v_mfma_f64_16x16x4f64
// tiles.  The taps are staged once per block in LDS; the pre-weight is fused
// into the A-operand staging and the diagonal term and post-scale into the store.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctlib {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int RT = 64;       // rows and output columns per block
constexpr int RKC = 32;      // contraction indices staged in LDS per step
constexpr int RPAD = 1;      // LDS row padding (doubles): 16 rows of one k land in distinct banks

// The RL taps vanish at even offsets other than 0 (h(n) = 0 for even n != 0), so output
// column j only needs the inputs m of the other parity, plus the diagonal h(0) * a[j].  For
// output parity p (j = 2 jj + p) and input parity q = 1 - p (m = 2 mm + q):
//   out[r][j] = sum_mm a[r][2 mm + q] * h(2 (mm - jj) + q - p)  +  h(0) a[r][j],
// two dense GEMMs of [rows x D/2] x [D/2 x D/2] instead of one [rows x D] x [D x D] with half
// its entries zero: half the MFMA work.
//
// Block (x, y, z = p): rows y*64 .. +63, output columns jj = x*64 .. +63 of parity p; each wave
// owns 16 rows x 64 columns (four 16x16 accumulators).  The pre-weighted A operand (fp64) is
// staged through LDS in [64 rows][32 mm] tiles read row-contiguously from global memory; the odd
// taps h(2 t + q - p), t = mm - jj, sit in LDS for the whole block.  Operand layout of
// v_mfma_f64_16x16x4f64 (checked on gfx950): A[i][k] and B[k][j] are held by lane i + 16k (resp.
// j + 16k); accumulator element r of lane l is D[l/16 + 4r][l%16].
template <typename TIn, typename TOut>
__global__ __launch_bounds__(256) void k_ramp(const TIn* __restrict__ in, TOut* __restrict__ out, long rows, int D,
                                             const double* __restrict__ taps, const double* __restrict__ pre,
                                             const double* __restrict__ post, double scale) {
    extern __shared__ double lds[];
    const int p = blockIdx.z, q = 1 - p;
    const int Dp = (D - p + 1) / 2;  // output columns of parity p
    const int Dq = (D - q + 1) / 2;  // inputs of parity q
    const int ntap = Dp + Dq - 1;    // t = mm - jj in [-(Dp-1), Dq-1]
    double* sh_t = lds;              // [ntap]: h(2 t + q - p) at t + Dp - 1
    double* sh_a = lds + ((ntap + 1) & ~1);  // [RT][RKC + RPAD]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < ntap; i += 256) sh_t[i] = taps[2 * (i - (Dp - 1)) + q - p + D - 1];
    const long r0 = (long)blockIdx.y * RT;
    const int jj0 = blockIdx.x * RT;
    const int li = lane & 15, kk = lane >> 4;
    dbl4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    // a[r0 + i][2 (m0 + c) + q] and pre[2 (m0 + c) + q] of the 64 x 32 tile at m0, 8 per thread
    // (row-contiguous), loaded one tile ahead so the loads overlap the previous tile's MFMAs
    constexpr int PER = RT * RKC / 256;
    TIn av[PER];
    double pv[PER];
    auto load_tile = [&](int m0) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u;
            const int i = e / RKC, c = e % RKC;
            const long r = r0 + i;
            const int mm = m0 + c, m = 2 * mm + q;
            const bool ok = r < rows && mm < Dq;
            av[u] = ok ? in[r * D + m] : TIn(0);
            pv[u] = ok && pre ? pre[m] : 1.0;
        }
    };
    load_tile(0);
    for (int m0 = 0; m0 < Dq; m0 += RKC) {
        __syncthreads();  // previous tile consumed (and, the first time, the taps written)
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int e = tid + 256 * u;
            sh_a[(e / RKC) * (RKC + RPAD) + e % RKC] = (double)av[u] * pv[u];
        }
        __syncthreads();
        if (m0 + RKC < Dq) load_tile(m0 + RKC);
        const double* ar = sh_a + (wave * 16 + li) * (RKC + RPAD);
#pragma unroll
        for (int c0 = 0; c0 < RKC; c0 += 4) {
            const int c = c0 + kk, mm = m0 + c;
            const double a = ar[c];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int jj = jj0 + 16 * t + li;
                // columns past Dp or inputs past Dq multiply a zero A entry or are never stored
                const int ti = mm - jj + Dp - 1;
                const double bv = (ti >= 0 && ti < ntap) ? sh_t[ti] : 0.0;
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
            }
        }
    }
    const double h0 = taps[D - 1];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int jj = jj0 + 16 * t + li;
        if (jj >= Dp) continue;
        const int col = 2 * jj + p;
        const double f = post ? post[col] * scale : scale;
        const double pw = pre ? pre[col] : 1.0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const long row = r0 + wave * 16 + kk + 4 * r;
            if (row < rows) out[row * D + col] = (TOut)((acc[t][r] + h0 * ((double)in[row * D + col] * pw)) * f);
        }
    }
}

template <typename TIn, typename TOut>
hipError_t ramp_t(const DevPlan& P, const void* in, void* out, long rows, const double* pre, const double* post,
                  double scale, hipStream_t s) {
    const int D = P.D;
    const int Dp0 = (D + 1) / 2;  // output columns of parity 0 (the larger count)
    dim3 grid((Dp0 + RT - 1) / RT, (unsigned)((rows + RT - 1) / RT), 2);
    const size_t lds = (size_t)(D + 1) * sizeof(double) + (size_t)RT * (RKC + RPAD) * sizeof(double);
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
