// kNN graph weights of the MAGIC network (`laplacian`, laplacian_cuda_kernel.cu:46-128)
// on gfx950, without the N x N distance matrix the reference materialises.
//
//   k_lap_norm     ||x_i||^2, sequential over the dimension (as :9-22)
//   k_lap_knn      one query point per thread, its coordinates in registers; the
//                  candidates of one slice of the point set stream through LDS in
//                  tiles of 64 points, each distance ||x_q||^2 + ||x_c||^2 - 2 x_q.x_c
//                  is formed with a sequential FMA chain and offered to the thread's
//                  sorted list of the KMAX smallest (value, column) pairs held in
//                  registers.  The grid is (query blocks) x (candidate slices), so a
//                  few thousand points still fill the 256 CUs.
//   k_lap_merge    per query: the slices' sorted lists -> the k smallest
//   k_lap_hist / k_lap_pick
//                  lower median of the N*k selected distances (torch.median, :115):
//                  8-bit radix select, each pass a grid-wide histogram (LDS, then
//                  global atomics) and a one-block digit pick
//   k_lap_weights  exp(-v / median), row-normalised, stored as float32 (:116-126)
//
// Dimensions or k beyond the register budget of k_lap_knn take the general path:
//   k_lap_dist     distances of a chunk of query rows to every point, 64 x 64 tiles with
//                  the dimension staged through LDS 16 at a time (the same sequential FMA
//                  chain per pair as k_lap_knn, so both paths give identical distances)
//   k_lap_select   one block per query row: the k-th smallest distance key by 8-bit radix
//                  select, the k smallest (value, column) pairs gathered (equal values by
//                  ascending column) and sorted by a bitonic sort in LDS
#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <type_traits>

#include "../../include/ctlib_abi.h"
#include "kernels.h"

namespace ctlib {
namespace {

constexpr int LAP_NT = 256;   // threads (query points) per k_lap_knn block
constexpr int LAP_TC = 64;    // candidate points per LDS tile
constexpr int LAP_GRID = 2048;  // blocks the (query, slice) grid aims for

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

// Offer (v, c) to a sorted register list of KMAX entries (static indices only).
template <typename T, int KMAX>
__device__ __forceinline__ void lap_offer(T (&lv)[KMAX], int (&li)[KMAX], T v, int c) {
    if (!lap_before(v, c, lv[KMAX - 1], li[KMAX - 1])) return;
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

constexpr int LAP_SB = 16;  // candidates per deferred-insertion batch

// Block (bx, by): queries bx*256 .. +255 against candidate slice by (tiles
// [by*tps, (by+1)*tps) of 64 points).  Writes the slice's sorted KMAX-list of
// every query to part[by][q][0..KMAX).  Insertions are deferred: per batch of
// 16 candidates each thread appends the ones below its current k-th entry to
// its own LDS column, then the wave inserts while any lane has one left, so a
// wave runs as many insertions as its busiest lane accepted instead of one for
// every candidate that some lane accepted.
template <typename T, int DMAX, int KMAX>
__global__ __launch_bounds__(256) void k_lap_knn(const T* __restrict__ x, const T* __restrict__ norm, int n, int d,
                                                 int tps, T* __restrict__ pv, int* __restrict__ pc) {
    __shared__ T ct[LAP_TC][DMAX];
    __shared__ T cn[LAP_TC];
    __shared__ T bv[LAP_SB][LAP_NT];  // accepted (value, column) per thread
    __shared__ int bc[LAP_SB][LAP_NT];
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
    const int cbeg = blockIdx.y * tps * LAP_TC;
    const int cend = min(n, cbeg + tps * LAP_TC);
    for (int c0 = cbeg; c0 < cend; c0 += LAP_TC) {
        const int nc = cend - c0 < LAP_TC ? cend - c0 : LAP_TC;
        __syncthreads();  // the previous tile is no longer read
        for (int e = tid; e < LAP_TC * DMAX; e += LAP_NT) {
            const int ci = e / DMAX, j = e % DMAX;
            ct[ci][j] = (ci < nc && j < d) ? x[(int64_t)(c0 + ci) * d + j] : T(0);
        }
        if (tid < LAP_TC) cn[tid] = tid < nc ? norm[c0 + tid] : T(0);
        __syncthreads();
        if (!active) continue;
        for (int cb = 0; cb < nc; cb += LAP_SB) {
            const int ce = nc < cb + LAP_SB ? nc : cb + LAP_SB;
            const T tv = lv[KMAX - 1];
            const int tc = li[KMAX - 1];
            int cnt = 0;
            for (int ci = cb; ci < ce; ci++) {
                T dot = 0;
#pragma unroll
                for (int j = 0; j < DMAX; j++) dot = fma(qv[j], ct[ci][j], dot);
                const T v = fma(T(-2), dot, nq + cn[ci]);
                if (lap_before(v, c0 + ci, tv, tc)) {
                    bv[cnt][tid] = v;
                    bc[cnt][tid] = c0 + ci;
                    cnt++;
                }
            }
            for (int i = 0; __ballot(i < cnt) != 0; i++)
                if (i < cnt) lap_offer<T, KMAX>(lv, li, bv[i][tid], bc[i][tid]);
        }
    }
    if (!active) return;
    const size_t o = ((size_t)blockIdx.y * n + q) * KMAX;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        pv[o + j] = lv[j];
        pc[o + j] = li[j];
    }
}

// One thread per query: the k smallest of the slices' lists, in order.
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_lap_merge(const T* __restrict__ pv, const int* __restrict__ pc, int n,
                                                   int slices, int k, int64_t* __restrict__ cols, T* __restrict__ vals) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    T lv[KMAX];
    int li[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        lv[j] = __builtin_inf();
        li[j] = 0x7fffffff;
    }
    for (int sl = 0; sl < slices; sl++) {
        const size_t o = ((size_t)sl * n + q) * KMAX;
        for (int j = 0; j < KMAX; j++) {
            const T v = pv[o + j];
            const int c = pc[o + j];
            if (!lap_before(v, c, lv[KMAX - 1], li[KMAX - 1])) break;  // the slice's list is sorted
            lap_offer<T, KMAX>(lv, li, v, c);
        }
    }
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

// Radix-select state in device memory: histogram of the current pass, the key
// prefix found so far and the rank still to skip within it.
struct LapSel {
    unsigned int hist[256];
    unsigned long long prefix;
    unsigned long long rank;
};

__global__ void k_lap_sel_init(LapSel* st, int64_t m) {
    const int t = threadIdx.x;
    st->hist[t] = 0u;
    if (t == 0) {
        st->prefix = 0ull;
        st->rank = (unsigned long long)((m - 1) / 2);  // torch.median: the lower middle element
    }
}

// Histogram of digit (key >> shift) & 255 over the values whose higher bits equal the prefix.
template <typename T>
__global__ __launch_bounds__(256) void k_lap_hist(const T* __restrict__ v, int64_t m, int shift, LapSel* st) {
    using K = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    constexpr int BITS = sizeof(T) * 8;
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const K prefix = (K)st->prefix;
    const K hmask = shift + 8 >= BITS ? K(0) : (~K(0) << (shift + 8));
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const K key = lap_key(v[i]);
        if ((key & hmask) == prefix) atomicAdd(&h[(key >> shift) & 255], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
}

// Digit of the rank-th value in this pass; clears the histogram for the next pass.
template <typename T>
__global__ void k_lap_pick(LapSel* st, int shift, T* __restrict__ out) {
    if (threadIdx.x == 0) {
        unsigned long long r = st->rank, acc = 0;
        int dig = 0;
        for (; dig < 255; dig++) {
            if (acc + st->hist[dig] > r) break;
            acc += st->hist[dig];
        }
        st->rank = r - acc;
        st->prefix |= (unsigned long long)dig << shift;
        if (shift == 0) {
            using K = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
            out[0] = lap_unkey((K)st->prefix);
        }
    }
    __syncthreads();
    st->hist[threadIdx.x] = 0u;
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

// ---- general path -------------------------------------------------------------------------

constexpr int LG_T = 64;     // query rows x candidate columns per k_lap_dist tile
constexpr int LG_KD = 16;    // dimensions staged in LDS per step
constexpr int LG_KMAX = 4096;  // largest k of the general path (sort buffer in LDS)

// dist[r][c] = ||x_q||^2 + ||x_c||^2 - 2 x_q . x_c for query q = q0 + r (r < nq) and every
// point c < n; each thread owns a 4 x 4 block of the 64 x 64 tile.
template <typename T>
__global__ __launch_bounds__(256) void k_lap_dist(const T* __restrict__ x, const T* __restrict__ norm, int n, int d,
                                                  int q0, int nq, T* __restrict__ dist) {
    constexpr int PAD = 16 / sizeof(T);  // keeps every LDS row 16-byte aligned
    __shared__ __align__(16) T qs[LG_KD][LG_T + PAD];
    __shared__ __align__(16) T cs[LG_KD][LG_T + PAD];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int qb = q0 + blockIdx.y * LG_T;  // first query of the tile
    const int cb = blockIdx.x * LG_T;       // first candidate of the tile
    const int qend = q0 + nq;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int l = 0; l < 4; l++) acc[i][l] = T(0);
    for (int j0 = 0; j0 < d; j0 += LG_KD) {
        for (int e = tid; e < LG_T * LG_KD; e += 256) {
            const int r = e / LG_KD, j = e % LG_KD;
            const int qi = qb + r, ci = cb + r;
            qs[j][r] = (qi < qend && j0 + j < d) ? x[(int64_t)qi * d + j0 + j] : T(0);
            cs[j][r] = (ci < n && j0 + j < d) ? x[(int64_t)ci * d + j0 + j] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < LG_KD; j++) {
            T a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                a[i] = qs[j][ty * 4 + i];
                b[i] = cs[j][tx * 4 + i];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int l = 0; l < 4; l++) acc[i][l] = fma(a[i], b[l], acc[i][l]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int qi = qb + ty * 4 + i;
        if (qi >= qend) continue;
        const T nqv = norm[qi];
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const int ci = cb + tx * 4 + l;
            if (ci < n) dist[(int64_t)(qi - q0) * n + ci] = fma(T(-2), acc[i][l], nqv + norm[ci]);
        }
    }
}

// Row r of the chunk (query q0 + r): the k smallest (value, column) pairs of dist[r][0..n) in
// ascending order, ties by column, into cols[q][0..k) and sel[q][0..k).  P: power of two >= k
// (the LDS sort buffer, P keys and P columns).
template <typename T>
__global__ __launch_bounds__(256) void k_lap_select(const T* __restrict__ dist, int n, int k, int q0, int P,
                                                    int64_t* __restrict__ cols, T* __restrict__ sel) {
    using K = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    constexpr int BITS = sizeof(T) * 8;
    extern __shared__ __align__(16) unsigned char lsm[];
    K* sk = reinterpret_cast<K*>(lsm);
    int* sc = reinterpret_cast<int*>(sk + P);
    __shared__ unsigned int hist[256];
    __shared__ unsigned int wsum[4];
    __shared__ unsigned int s_dig, s_rank, s_lt;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = blockIdx.x;
    const T* dr = dist + (int64_t)r * n;

    // 1. key of rank k-1 (0-based) by 8-bit radix select over the row
    K prefix = 0, hmask = 0;
    unsigned int rank = (unsigned int)(k - 1);
    for (int shift = BITS - 8; shift >= 0; shift -= 8) {
        hist[tid] = 0u;
        __syncthreads();
        for (int c = tid; c < n; c += 256) {
            const K key = lap_key(dr[c]);
            if ((key & hmask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            unsigned int accu = 0;
            int dig = 0;
            for (; dig < 255; dig++) {
                if (accu + hist[dig] > rank) break;
                accu += hist[dig];
            }
            s_dig = (unsigned int)dig;
            s_rank = rank - accu;
        }
        __syncthreads();
        prefix |= (K)s_dig << shift;
        hmask |= (K)255 << shift;
        rank = s_rank;
    }
    // rank = how many values equal to the k-th one precede it, so k-1-rank values are smaller
    const K kth = prefix;
    const int nlt = k - 1 - (int)rank;
    const int neq = (int)rank + 1;

    // 2. gather: the smaller values in any order, the equal ones by ascending column
    if (tid == 0) s_lt = 0u;
    for (int i = tid; i < P; i += 256) {
        sk[i] = ~K(0);
        sc[i] = 0x7fffffff;
    }
    __syncthreads();
    int eq_base = 0;
    for (int c0 = 0; c0 < n; c0 += 256) {
        const int c = c0 + tid;
        const K key = c < n ? lap_key(dr[c]) : ~K(0);
        const bool lt = c < n && key < kth;
        const bool eq = c < n && key == kth;
        if (lt) {
            const unsigned int pos = atomicAdd(&s_lt, 1u);
            sk[pos] = key;
            sc[pos] = c;
        }
        const unsigned long long bal = __ballot(eq);
        const int below = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = (unsigned int)__popcll(bal);
        __syncthreads();
        int before = eq_base;
        for (int w = 0; w < wave; w++) before += (int)wsum[w];
        const int pos = before + below;
        if (eq && pos < neq) {
            sk[nlt + pos] = key;
            sc[nlt + pos] = c;
        }
        eq_base += (int)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
        const unsigned int lt_now = s_lt;  // every atomic of this step precedes the barrier above
        __syncthreads();
        if (eq_base >= neq && (int)lt_now >= nlt) break;  // uniform: all threads read the same counts
    }

    // 3. bitonic sort of the P entries by (key, column)
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < P / 2; i += 256) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const K ka = sk[lo], kb = sk[hi];
                const int ca = sc[lo], cb = sc[hi];
                const bool gt = ka > kb || (ka == kb && ca > cb);
                if (gt == up) {
                    sk[lo] = kb;
                    sk[hi] = ka;
                    sc[lo] = cb;
                    sc[hi] = ca;
                }
            }
            __syncthreads();
        }
    }
    const int64_t q = (int64_t)q0 + r;
    for (int j = tid; j < k; j += 256) {
        cols[q * k + j] = sc[j];
        sel[q * k + j] = lap_unkey(sk[j]);
    }
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

// Candidate slices: enough (query block, slice) pairs for the chip, at least one tile each.
int lap_slices(int64_t n) {
    const int64_t qb = (n + LAP_NT - 1) / LAP_NT;
    const int64_t tiles = (n + LAP_TC - 1) / LAP_TC;
    int64_t s = (LAP_GRID + qb - 1) / qb;
    if (s > tiles) s = tiles;
    if (s > 64) s = 64;
    return (int)(s < 1 ? 1 : s);
}

size_t lap_align(size_t b) { return (b + 255) & ~(size_t)255; }

// Query rows per chunk of the general path: distance rows of at most 128 MiB, a multiple of 64.
int lap_chunk_rows(int64_t n, int dtype) {
    const size_t es = dtype == CTLIB_F64 ? 8 : 4;
    int64_t q = ((int64_t)128 << 20) / ((int64_t)n * (int64_t)es);
    q = q / LG_T * LG_T;
    if (q < LG_T) q = LG_T;
    if (q > n) q = n;
    return (int)q;
}

struct LapWs {
    size_t norm, sel, part_v, part_c, state, median, total;
};

// kmax > 0: register path (slices' partial lists in part_v / part_c); kmax == 0: general path
// (part_v holds the distance rows of one chunk).
LapWs lap_layout(int64_t n, int k, int dtype, int kmax) {
    const size_t es = dtype == CTLIB_F64 ? 8 : 4;
    const size_t parts = kmax ? (size_t)lap_slices(n) * n * kmax : (size_t)lap_chunk_rows(n, dtype) * n;
    LapWs L;
    size_t off = 0;
    L.norm = off; off += lap_align(n * es);
    L.sel = off; off += lap_align((size_t)n * k * es);
    L.part_v = off; off += lap_align(parts * es);
    L.part_c = off; off += lap_align(kmax ? parts * sizeof(int) : 0);
    L.state = off; off += lap_align(sizeof(LapSel));
    L.median = off; off += lap_align(es);
    L.total = off;
    return L;
}

int lap_pow2(int k) {
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

// General path: distances chunk by chunk, then the per-row selection.
template <typename T>
hipError_t knn_general(const void* x, const void* norm, int n, int d, int k, int64_t* cols, char* w, const LapWs& L,
                       hipStream_t s) {
    const int rows = lap_chunk_rows(n, sizeof(T) == 8 ? CTLIB_F64 : CTLIB_F32);
    const int P = lap_pow2(k);
    const size_t lds = (size_t)P * (sizeof(T) + sizeof(int));
    T* dist = reinterpret_cast<T*>(w + L.part_v);
    for (int q0 = 0; q0 < n; q0 += rows) {
        const int nq = n - q0 < rows ? n - q0 : rows;
        const dim3 g((n + LG_T - 1) / LG_T, (nq + LG_T - 1) / LG_T);
        hipLaunchKernelGGL(k_lap_dist<T>, g, dim3(256), 0, s, (const T*)x, (const T*)norm, n, d, q0, nq, dist);
        hipLaunchKernelGGL(k_lap_select<T>, dim3(nq), dim3(256), lds, s, (const T*)dist, n, k, q0, P, cols,
                           reinterpret_cast<T*>(w + L.sel));
    }
    return hipGetLastError();
}

template <typename T, int DMAX, int KMAX>
hipError_t knn_launch(const void* x, const void* norm, int n, int d, int k, int64_t* cols, char* w, const LapWs& L,
                      hipStream_t s) {
    const int slices = lap_slices(n);
    const int tiles = (n + LAP_TC - 1) / LAP_TC;
    const int tps = (tiles + slices - 1) / slices;
    const dim3 grid((n + LAP_NT - 1) / LAP_NT, slices);
    hipLaunchKernelGGL((k_lap_knn<T, DMAX, KMAX>), grid, dim3(LAP_NT), 0, s, (const T*)x, (const T*)norm, n, d, tps,
                       (T*)(w + L.part_v), (int*)(w + L.part_c));
    hipLaunchKernelGGL((k_lap_merge<T, KMAX>), dim3((n + 255) / 256), dim3(256), 0, s, (const T*)(w + L.part_v),
                       (const int*)(w + L.part_c), n, slices, k, cols, (T*)(w + L.sel));
    return hipGetLastError();
}

template <typename T, int DMAX>
hipError_t knn_k(const LapShape& sh, const void* x, const void* norm, int n, int d, int k, int64_t* cols, char* w,
                 const LapWs& L, hipStream_t s) {
    switch (sh.kmax) {
    case 8: return knn_launch<T, DMAX, 8>(x, norm, n, d, k, cols, w, L, s);
    case 16: return knn_launch<T, DMAX, 16>(x, norm, n, d, k, cols, w, L, s);
    case 32:
        if constexpr (DMAX * (sizeof(T) / 4) + 32 * (sizeof(T) / 4 + 1) <= 192)
            return knn_launch<T, DMAX, 32>(x, norm, n, d, k, cols, w, L, s);
        return hipErrorInvalidValue;
    default:
        if constexpr (DMAX * (sizeof(T) / 4) + 64 * (sizeof(T) / 4 + 1) <= 192)
            return knn_launch<T, DMAX, 64>(x, norm, n, d, k, cols, w, L, s);
        return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t run(const LapShape& sh, const void* x, int64_t n, int d, int k, int64_t* cols, float* vals, char* w,
               const LapWs& L, hipStream_t s) {
    const dim3 g1((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_lap_norm<T>, g1, dim3(256), 0, s, (const T*)x, n, d, (T*)(w + L.norm));
    hipError_t e;
    switch (sh.kmax ? sh.dmax : 0) {
    case 0: e = knn_general<T>(x, w + L.norm, (int)n, d, k, cols, w, L, s); break;
    case 16: e = knn_k<T, 16>(sh, x, w + L.norm, (int)n, d, k, cols, w, L, s); break;
    case 32: e = knn_k<T, 32>(sh, x, w + L.norm, (int)n, d, k, cols, w, L, s); break;
    case 64: e = knn_k<T, 64>(sh, x, w + L.norm, (int)n, d, k, cols, w, L, s); break;
    default: e = knn_k<T, 128>(sh, x, w + L.norm, (int)n, d, k, cols, w, L, s); break;
    }
    if (e != hipSuccess) return e;
    const int64_t m = n * k;
    LapSel* st = reinterpret_cast<LapSel*>(w + L.state);
    T* med = reinterpret_cast<T*>(w + L.median);
    const T* sel = reinterpret_cast<const T*>(w + L.sel);
    hipLaunchKernelGGL(k_lap_sel_init, dim3(1), dim3(256), 0, s, st, m);
    const int hb = (int)std::min<int64_t>(1024, (m + 255) / 256);
    for (int shift = (int)sizeof(T) * 8 - 8; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(k_lap_hist<T>, dim3(hb), dim3(256), 0, s, sel, m, shift, st);
        hipLaunchKernelGGL(k_lap_pick<T>, dim3(1), dim3(256), 0, s, st, shift, med);
    }
    hipLaunchKernelGGL(k_lap_weights<T>, g1, dim3(256), 0, s, sel, (const T*)med, n, k, vals);
    return hipGetLastError();
}

}  // namespace

// Register path when (d, k) fit its budget, else the general path (any d, k <= LG_KMAX).
bool lap_pick(int d, int k, int dtype, LapShape* sh) {
    if (lap_shape(d, k, dtype, sh)) return true;
    sh->dmax = 0;
    sh->kmax = 0;
    return d >= 1 && k >= 1 && k <= LG_KMAX;
}

bool laplacian_supported(int d, int k, int dtype) {
    LapShape sh;
    return lap_pick(d, k, dtype, &sh);
}

size_t laplacian_ws_bytes(int64_t n, int d, int k, int dtype) {
    LapShape sh;
    if (!lap_pick(d, k, dtype, &sh)) return 0;
    return lap_layout(n, k, dtype, sh.kmax).total;
}

hipError_t launch_laplacian(const void* x, int64_t n, int d, int k, int dtype, int64_t* cols, float* vals, void* ws,
                            hipStream_t s) {
    LapShape sh;
    if (!lap_pick(d, k, dtype, &sh)) return hipErrorInvalidValue;
    const LapWs L = lap_layout(n, k, dtype, sh.kmax);
    char* w = static_cast<char*>(ws);
    if (dtype == CTLIB_F32) return run<float>(sh, x, n, d, k, cols, vals, w, L, s);
    return run<double>(sh, x, n, d, k, cols, vals, w, L, s);
}

}  // namespace ctlib
