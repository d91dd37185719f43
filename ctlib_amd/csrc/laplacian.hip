// kNN graph weights of the MAGIC network (`laplacian`, laplacian_cuda_kernel.cu:46-128)
// on gfx950, without the N x N distance matrix the reference materialises.
//
//   k_lap_norm     ||x_i||^2, sequential over the dimension (as :9-22)
//   k_lap_knn      one query point per thread, its coordinates in registers; the
//                  candidates stream through LDS in tiles of 64 points, each
//                  distance ||x_q||^2 + ||x_c||^2 - 2 x_q.x_c is formed with a
//                  sequential FMA chain and offered to the thread's sorted list
//                  of the KMAX smallest (value, column) pairs held in registers
//   k_lap_median   lower median of the N*k selected distances (torch.median, :115)
//                  by an 8-bit radix select in one block (LDS histograms)
//   k_lap_weights  exp(-v / median), row-normalised, stored as float32 (:116-126)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../include/ctlib_abi.h"
#include "kernels.h"

namespace ctlib {
namespace {

constexpr int LAP_NT = 256;  // threads (query points) per k_lap_knn block
constexpr int LAP_TC = 64;   // candidate points per LDS tile

template <typename T>
__global__ __launch_bounds__(256) void k_lap_norm(const T* __restrict__ x, int64_t n, int d, T* __restrict__ norm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T* r = x + i * d;
    T s = 0;
    for (int j = 0; j < d; j++) s = fma(r[j], r[j], s);
    norm[i] = s;
}

// (value, column) order of the lists: ascending value, ties by ascending column.
template <typename T>
__device__ __forceinline__ bool lap_before(T v, int c, T w, int e) {
    return v < w || (v == w && c < e);
}

template <typename T, int DMAX, int KMAX>
__global__ __launch_bounds__(256) void k_lap_knn(const T* __restrict__ x, const T* __restrict__ norm, int n, int d,
                                                 int k, int64_t* __restrict__ cols, T* __restrict__ vals) {
    __shared__ T ct[LAP_TC][DMAX];
    __shared__ T cn[LAP_TC];
    const int tid = threadIdx.x;
    const int q = blockIdx.x * LAP_NT + tid;
    const bool active = q < n;
    T qv[DMAX];
#pragma unroll
    for (int j = 0; j < DMAX; j++) qv[j] = (active && j < d) ? x[(int64_t)q * d + j] : T(0);
    const T nq = active ? norm[q] : T(0);
    T lv[KMAX];
    int li[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        lv[j] = __builtin_inf();
        li[j] = 0x7fffffff;
    }
    for (int c0 = 0; c0 < n; c0 += LAP_TC) {
        const int nc = n - c0 < LAP_TC ? n - c0 : LAP_TC;
        __syncthreads();  // the previous tile is no longer read
        for (int e = tid; e < LAP_TC * DMAX; e += LAP_NT) {
            const int ci = e / DMAX, j = e % DMAX;
            ct[ci][j] = (ci < nc && j < d) ? x[(int64_t)(c0 + ci) * d + j] : T(0);
        }
        if (tid < LAP_TC) cn[tid] = tid < nc ? norm[c0 + tid] : T(0);
        __syncthreads();
        if (!active) continue;
        for (int ci = 0; ci < nc; ci++) {
            T dot = 0;
#pragma unroll
            for (int j = 0; j < DMAX; j++) dot = fma(qv[j], ct[ci][j], dot);
            T v = fma(T(-2), dot, nq + cn[ci]);
            int c = c0 + ci;
            if (lap_before(v, c, lv[KMAX - 1], li[KMAX - 1])) {
                // insertion into the sorted register list (static indices only)
#pragma unroll
                for (int j = 0; j < KMAX; j++) {
                    const bool lt = lap_before(v, c, lv[j], li[j]);
                    const T tv = lt ? lv[j] : v;
                    const int tc = lt ? li[j] : c;
                    lv[j] = lt ? v : lv[j];
                    li[j] = lt ? c : li[j];
                    v = tv;
                    c = tc;
                }
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j < k) {
            cols[(int64_t)q * k + j] = li[j];
            vals[(int64_t)q * k + j] = lv[j];
        }
    }
}

// Order-preserving unsigned keys of floating-point values.
__device__ __forceinline__ uint32_t lap_key(float v) {
    uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t lap_key(double v) {
    uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ float lap_unkey(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ double lap_unkey(uint64_t u) {
    return __longlong_as_double((long long)((u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u));
}

// Value of rank (m - 1) / 2 among m values (torch.median's lower median): radix
// select, 8 bits per pass from the top, one 1024-thread block.
template <typename T>
__global__ __launch_bounds__(1024) void k_lap_median(const T* __restrict__ v, int64_t m, T* __restrict__ out) {
    using K = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    constexpr int BITS = sizeof(T) * 8;
    __shared__ unsigned int hist[256];
    __shared__ K s_prefix;
    __shared__ unsigned long long s_rank;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_prefix = 0;
        s_rank = (unsigned long long)((m - 1) / 2);
    }
    for (int shift = BITS - 8; shift >= 0; shift -= 8) {
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        const K prefix = s_prefix;
        const K hmask = shift + 8 >= BITS ? K(0) : (~K(0) << (shift + 8));
        for (int64_t i = tid; i < m; i += blockDim.x) {
            const K key = lap_key(v[i]);
            if ((key & hmask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long r = s_rank, acc = 0;
            int dig = 0;
            for (; dig < 255; dig++) {
                if (acc + hist[dig] > r) break;
                acc += hist[dig];
            }
            s_rank = r - acc;
            s_prefix = prefix | ((K)dig << shift);
        }
        __syncthreads();
    }
    if (tid == 0) out[0] = lap_unkey(s_prefix);
}

// One thread per point: w_j = exp(-v_j / median) / sum_j exp(-v_j / median), as float32.
template <typename T>
__global__ __launch_bounds__(256) void k_lap_weights(const T* __restrict__ vals, const T* __restrict__ median,
                                                     int64_t n, int k, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T md = median[0];
    const T* r = vals + i * k;
    T s = 0;
    for (int j = 0; j < k; j++) s += exp(-(r[j] / md));
    for (int j = 0; j < k; j++) out[i * k + j] = (float)(exp(-(r[j] / md)) / s);
}

struct LapShape {
    int dmax, kmax;
};

// Register budget of k_lap_knn: DMAX coordinates plus KMAX (value, column) pairs.
bool lap_shape(int d, int k, int dtype, LapShape* s) {
    static const int dm[] = {16, 32, 64, 128};
    static const int km[] = {8, 16, 32, 64};
    s->dmax = 0;
    for (int v : dm)
        if (d <= v) { s->dmax = v; break; }
    s->kmax = 0;
    for (int v : km)
        if (k <= v) { s->kmax = v; break; }
    if (!s->dmax || !s->kmax) return false;
    const int words = dtype == CTLIB_F64 ? 2 : 1;
    return s->dmax * words + s->kmax * (words + 1) <= 192;
}

template <typename T, int DMAX, int KMAX>
hipError_t knn_launch(const void* x, const void* norm, int n, int d, int k, int64_t* cols, void* vals, hipStream_t s) {
    hipLaunchKernelGGL((k_lap_knn<T, DMAX, KMAX>), dim3((n + LAP_NT - 1) / LAP_NT), dim3(LAP_NT), 0, s, (const T*)x,
                       (const T*)norm, n, d, k, cols, (T*)vals);
    return hipGetLastError();
}

template <typename T, int DMAX>
hipError_t knn_k(const LapShape& sh, const void* x, const void* norm, int n, int d, int k, int64_t* cols, void* vals,
                 hipStream_t s) {
    switch (sh.kmax) {
    case 8: return knn_launch<T, DMAX, 8>(x, norm, n, d, k, cols, vals, s);
    case 16: return knn_launch<T, DMAX, 16>(x, norm, n, d, k, cols, vals, s);
    case 32:
        if constexpr (DMAX * (sizeof(T) / 4) + 32 * (sizeof(T) / 4 + 1) <= 192)
            return knn_launch<T, DMAX, 32>(x, norm, n, d, k, cols, vals, s);
        return hipErrorInvalidValue;
    default:
        if constexpr (DMAX * (sizeof(T) / 4) + 64 * (sizeof(T) / 4 + 1) <= 192)
            return knn_launch<T, DMAX, 64>(x, norm, n, d, k, cols, vals, s);
        return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t knn(const LapShape& sh, const void* x, const void* norm, int n, int d, int k, int64_t* cols, void* vals,
               hipStream_t s) {
    switch (sh.dmax) {
    case 16: return knn_k<T, 16>(sh, x, norm, n, d, k, cols, vals, s);
    case 32: return knn_k<T, 32>(sh, x, norm, n, d, k, cols, vals, s);
    case 64: return knn_k<T, 64>(sh, x, norm, n, d, k, cols, vals, s);
    default: return knn_k<T, 128>(sh, x, norm, n, d, k, cols, vals, s);
    }
}

size_t lap_align(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

bool laplacian_supported(int d, int k, int dtype) {
    LapShape sh;
    return lap_shape(d, k, dtype, &sh);
}

size_t laplacian_ws_bytes(int64_t n, int k, int dtype) {
    const size_t es = dtype == CTLIB_F64 ? 8 : 4;
    return lap_align(n * es) + lap_align((size_t)n * k * es) + lap_align(es);
}

hipError_t launch_laplacian(const void* x, int64_t n, int d, int k, int dtype, int64_t* cols, float* vals, void* ws,
                            hipStream_t s) {
    LapShape sh;
    if (!lap_shape(d, k, dtype, &sh)) return hipErrorInvalidValue;
    const size_t es = dtype == CTLIB_F64 ? 8 : 4;
    char* w = static_cast<char*>(ws);
    void* norm = w;
    void* sel = w + lap_align(n * es);
    void* med = w + lap_align(n * es) + lap_align((size_t)n * k * es);
    const dim3 g1((unsigned)((n + 255) / 256));
    hipError_t e;
    if (dtype == CTLIB_F32) {
        hipLaunchKernelGGL(k_lap_norm<float>, g1, dim3(256), 0, s, (const float*)x, n, d, (float*)norm);
        e = knn<float>(sh, x, norm, (int)n, d, k, cols, sel, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_lap_median<float>, dim3(1), dim3(1024), 0, s, (const float*)sel, n * k, (float*)med);
        hipLaunchKernelGGL(k_lap_weights<float>, g1, dim3(256), 0, s, (const float*)sel, (const float*)med, n, k, vals);
    } else {
        hipLaunchKernelGGL(k_lap_norm<double>, g1, dim3(256), 0, s, (const double*)x, n, d, (double*)norm);
        e = knn<double>(sh, x, norm, (int)n, d, k, cols, sel, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_lap_median<double>, dim3(1), dim3(1024), 0, s, (const double*)sel, n * k, (double*)med);
        hipLaunchKernelGGL(k_lap_weights<double>, g1, dim3(256), 0, s, (const double*)sel, (const double*)med, n, k,
                           vals);
    }
    return hipGetLastError();
}

}  // namespace ctlib
