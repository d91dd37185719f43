// Device-side building blocks of the distance-driven projector kernels
// (gfx950).  Each operator is written as a per-work-item block function so the
// stand-alone kernels (dd_kernels.hip) and the persistent heterogeneous batch
// kernel (batch.hip) run the same code.
//
// Every operator of the family (fan_ed_kernel.cu, fan_ea_kernel.cu,
// para_kernel.cu) is, per view and per image line, the overlap-weighted
// resampling of a piecewise-constant function between two sorted edge grids
// on the sweep axis: the D+1 detector-bin edges e_k and the n+1 pixel edges
// pi_p = a + s*p of the line.  The reference walks both grids with a
// data-dependent 4-way merge per thread and scatters with atomics.  Here both
// directions are evaluated branch-free through FP64 cumulative integrals:
//
//   line -> bin (projection, backprojection_t):
//     F_l(tau) = integral_0^tau line(t) dt  (tau in pixel units), stored as a
//     table {C_j, slope_j} so that F_l(tau) = C_j + tau*slope_j with
//     j = trunc(tau + 1) clamped to [0, n+1].  Each thread owns a few bin edges
//     and accumulates  S_k = sum_l s_l * F_l(tau_l(e_k));  the bin value is
//     post_k * (S_{k+1} - S_k).  One table per line is shared by G views.
//
//   bin -> pixel (projection_t, backprojection):
//     G_v(e) = integral of the per-bin values over the axis, stored per view as
//     {C_j, g_j} with G_v(e) = C_j + e*g_j and j found by a closed-form
//     (Moebius / arctangent / affine) axis -> bin map.  Each lane owns pixel
//     edges; projection_t accumulates T_p = sum_v G_v(pi_p) and the pixel value
//     is T_{p+1} - T_p; backprojection weights each per-view difference.
//
// Accumulation is FP64 throughout (FP64 FMA issues at the FP32 rate on CDNA4);
// only the image / sinogram storage is fp32 (or fp64).  No atomics: every
// output element is written once, so results are bitwise deterministic.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctlib {
namespace dd {

constexpr int NT = 256;      // threads per block (4 wave64)
constexpr int LC_WORDS = 8;  // doubles per LineCoef: a, s, A, B, w0, w1, w2, inv_s
enum { LC_A0 = 0, LC_S = 1, LC_A = 2, LC_B = 3, LC_W0 = 4, LC_W1 = 5, LC_W2 = 6, LC_INVS = 7 };
enum { GEO_ED = 0, GEO_EA = 1, GEO_PARA = 2, GEO_EAP = 3 };  // EAP: fan-EA, polynomial atan

// Blocks are dispatched round-robin over the 8 XCDs (block id % 8).  Give each
// XCD a contiguous range of work items so that blocks sharing an image or a
// sinogram (same batch element) share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks, int nwork) {
    const int per = nblocks >> 3;  // nblocks is a multiple of 8
    const int w = (bid & 7) * per + (bid >> 3);
    return w < nwork ? w : -1;
}

inline int grid_for(int nwork) { return ((nwork + 7) / 8) * 8; }

// 1/d for fp64 data: the fp32 hardware reciprocal (relative error ~1e-7) refined by one
// Newton step, r + r (1 - d r), to ~1e-14.  Used for squared source-pixel distances and
// lookup denominators, which valid geometries keep well inside the fp32 normal range.
__device__ __forceinline__ double rcp_refined(double d) {
    const double r = (double)__builtin_amdgcn_rcpf((float)d);
    return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}

// Index of a cumulative table from a float coordinate, clamped to [0, hi].
__device__ __forceinline__ int table_index(float q, float hi) { return (int)__builtin_amdgcn_fmed3f(q, 0.0f, hi); }

// The same from a double coordinate: v_cvt_i32_f64 (truncation toward zero)
// and one v_med3_i32 clamp to [0, hi] (written out because the compiler emits
// a max/min pair for it).  Coordinates of valid geometries stay far inside the
// int32 range.
__device__ __forceinline__ int table_index_d(double q, int hi) {
    int j = (int)q, r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(j), "v"(hi));
    return r;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int2 w = *reinterpret_cast<int2*>(&v);
    int2 r;
    r.x = __builtin_amdgcn_readlane(w.x, lane);
    r.y = __builtin_amdgcn_readlane(w.y, lane);
    return *reinterpret_cast<double*>(&r);
}

// Lane-shifted copy of a double through DPP (both 32-bit halves; lanes the
// control does not write read 0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
    const int2 w = *reinterpret_cast<int2*>(&v);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(0, w.x, CTRL, ROW_MASK, 0xf, true);
    r.y = __builtin_amdgcn_update_dpp(0, w.y, CTRL, ROW_MASK, 0xf, true);
    return *reinterpret_cast<double*>(&r);
}

// Lane i receives lane i+1's value (DPP wave_shl:1, no LDS round trip); lane 63 receives 0.
__device__ __forceinline__ double shfl_next_d(double v) { return dpp_d<0x130, 0xf>(v); }

// Lane i receives lane (i+1) mod 64's value (DPP wave_rol:1).
__device__ __forceinline__ double rot_next_d(double v) { return dpp_d<0x134, 0xf>(v); }

// Inclusive wave64 scan without LDS round trips: row_shr 1, 2, 4, 8 scan each
// 16-lane row, then row_bcast:15 adds row 0's total to row 1 and row 2's to
// row 3, and row_bcast:31 adds lane 31's running total to rows 2 and 3.
__device__ __forceinline__ double wave_scan_incl(double v, int) {
    v += dpp_d<0x111, 0xf>(v);  // row_shr:1
    v += dpp_d<0x112, 0xf>(v);  // row_shr:2
    v += dpp_d<0x114, 0xf>(v);  // row_shr:4
    v += dpp_d<0x118, 0xf>(v);  // row_shr:8
    v += dpp_d<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3
    v += dpp_d<0x143, 0xc>(v);  // row_bcast:31 -> rows 2, 3
    return v;
}

// ---------------------------------------------------------------------------
// 32x32 tile transposes for the Y-axis lines (columns read bottom to top)
// ---------------------------------------------------------------------------

// xT[c][i] = x[H-1-i][c] for the tile at (r0, c0) of one image.
template <typename T>
__device__ __forceinline__ void ylines_tile(const T* __restrict__ xb, T* __restrict__ tb, int H, int W, int r0, int c0,
                                            T (*tile)[33]) {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < H && c < W) tile[ty + k][tx] = xb[(size_t)r * W + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int c = c0 + ty + k;   // output line
        const int r = r0 + 31 - tx;  // row read by this lane
        if (c < W && r < H) tb[(size_t)c * H + (H - 1 - r)] = tile[31 - tx][ty + k];
    }
}

// out[r][c] = (hasX ? out[r][c] : 0) + (hasY ? yt[c][H-1-r] : 0) for the tile at (r0, c0).
template <typename T>
__device__ __forceinline__ void combine_tile(T* __restrict__ ob, const T* __restrict__ yb, int H, int W, int r0, int c0,
                                             int hasX, int hasY, T (*tile)[33]) {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (hasY) {
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            const int c = c0 + ty + k;
            const int r = r0 + 31 - tx;
            if (c < W && r < H) tile[ty + k][31 - tx] = yb[(size_t)c * H + (H - 1 - r)];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < H && c < W) {
            T v = hasX ? ob[(size_t)r * W + c] : T(0);
            if (hasY) v += tile[tx][ty + k];
            ob[(size_t)r * W + c] = v;
        }
    }
}

// out[r][c] = sum_k xs[k][r][c] (hasX) + sum_k ys[k][c][H-1-r] (hasY) over the nch
// partial images of a view-chunked bin -> pixel launch, for the tile at (r0, c0);
// partial k of the X (Y) images starts k*pstride elements further.
template <typename T>
__device__ __forceinline__ void combine_parts_tile(T* __restrict__ ob, const T* __restrict__ xs,
                                                   const T* __restrict__ ys, int nch, size_t pstride, int H, int W,
                                                   int r0, int c0, int hasX, int hasY, T (*tile)[33]) {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (hasY) {
#pragma unroll
        for (int k = 0; k < 32; k += 8) {
            const int c = c0 + ty + k;
            const int r = r0 + 31 - tx;
            T v = T(0);
            if (c < W && r < H)
                for (int ch = 0; ch < nch; ch++) v += ys[ch * pstride + (size_t)c * H + (H - 1 - r)];
            tile[ty + k][31 - tx] = v;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < H && c < W) {
            T v = T(0);
            if (hasX)
                for (int ch = 0; ch < nch; ch++) v += xs[ch * pstride + (size_t)r * W + c];
            if (hasY) v += tile[tx][ty + k];
            ob[(size_t)r * W + c] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Line -> bin (projection / backprojection_t)
// ---------------------------------------------------------------------------

// Entry of a line table: F(tau) = c + tau * s, 16 bytes (one ds_read_b128).
// A 12-byte form (fp32 slope, ds_read_b96) measured 60 % slower on gfx950:
// the 16-byte-strided b96 reads conflict in LDS banks.
template <typename T>
struct __attribute__((aligned(16))) LEnt {
    double c;
    double s;
};

__device__ __forceinline__ void lent_read(const void* p, double& c, double& s) {
    const double2 v = *reinterpret_cast<const double2*>(p);
    c = v.x;
    s = v.y;
}
__device__ __forceinline__ void lent_write(void* p, double c, double s) {
    *reinterpret_cast<double2*>(p) = make_double2(c, s);
}

// Bin values y_k = post_k (S_{k+1} - S_k) from the edge sums of a wave's VPW
// views, formed with wave shifts; RESID also writes the residual bin tables of
// projection_t, tabs[b][view] = cumulative {C, g} of g = (y - m) * post, and, when
// rowsq is not null, rowsq[b][view] = sum_k (y_k - m_k)^2 with y_k rounded to T.
// EDGEX: D == 64*NEM, and the sum of the last edge D of view v sits in lane v of Sx.
template <typename T, int VPW, int NEM, bool RESID, bool EDGEX = false>
__device__ __forceinline__ void fwd_store(const DevPlan& P, const double (&S)[VPW][NEM], const int (&vw)[VPW], int grp,
                                          int b, int mode, int lane, int wave, T* __restrict__ y,
                                          const double* __restrict__ post, const T* __restrict__ m,
                                          double2* __restrict__ tabs, double Sx = 0.0,
                                          double* __restrict__ rowsq = nullptr) {
    constexpr int G = 4 * VPW;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const int D = P.D;
    const int ne = (D + 1 + 63) / 64;  // edge slots in use (<= NEM)
#pragma unroll
    for (int v = 0; v < VPW; v++) {
        const int idx = grp * G + wave * VPW + v;
        if (idx >= nv) continue;  // wave-uniform
        const int view = vw[v];
        const bool rev = P.vp[view].rev;
        const size_t row = (size_t)b * P.V + view;
        const double* pr = post ? post + (size_t)view * D : nullptr;
        const double* ev = P.edges + (size_t)view * (D + 1);
        double2* tb = RESID ? tabs + row * (D + 2) : nullptr;
        double carry = 0.0, rsq = 0.0;
#pragma unroll
        for (int i = 0; i < NEM; i++) {
            if (i >= ne) continue;  // wave-uniform
            const int k = lane + 64 * i;
            double nxt = shfl_next_d(S[v][i]);
            double r0 = 0.0;
#pragma unroll
            for (int q = 0; q < NEM; q++)
                if (q == i + 1) r0 = readlane_d(S[v][q], 0);
            if (EDGEX && i == NEM - 1) r0 = readlane_d(Sx, v);
            if (lane == 63) nxt = r0;
            double yk = nxt - S[v][i];
            if (pr && k < D) yk *= pr[k];
            const int o = rev ? D - 1 - k : k;
            if (!RESID) {
                if (k < D) y[row * D + o] = (T)yk;
            } else {
                double gk = 0.0, ek = 0.0, wg = 0.0;
                if (k < D) {
                    if (y) y[row * D + o] = (T)yk;
                    const double mk = (double)m[row * D + o];
                    gk = (yk - mk) * pr[k];
                    const double rk = (double)(T)yk - mk;
                    rsq = __builtin_fma(rk, rk, rsq);
                    ek = ev[k];
                    wg = (ev[k + 1] - ek) * gk;
                }
                const double inc = wave_scan_incl(wg, lane);
                if (k < D) tb[k + 1] = make_double2(carry + (inc - wg) - ek * gk, gk);
                carry += readlane_d(inc, 63);
            }
        }
        if (RESID && lane == 0) {
            tb[0] = make_double2(0.0, 0.0);
            tb[D + 1] = make_double2(carry, 0.0);
        }
        if (RESID && rowsq) {
            const double tot = readlane_d(wave_scan_incl(rsq, lane), 63);
            if (lane == 0) rowsq[row] = tot;
        }
    }
}

// Builds NTAB line tables {C_j, slope_j}, j = 0..n+1, table t from val[t] (this
// thread's pixels p0 .. p0+m-1; already rounded to the storage precision of
// the slope).  Table t starts at tb + t*tabn; `ws` holds NTAB x 4 wave sums.
// One barrier inside, whatever NTAB.  FULL: every thread owns exactly M pixels
// (n == 256*M), so no pixel predicate is needed.
template <typename T, int M, int NTAB, bool FULL = false>
__device__ __forceinline__ void line_tables(const double (&val)[NTAB][M], int m, int p0, int n, LEnt<T>* __restrict__ tb,
                                            int tabn, double* __restrict__ ws, int tid, int lane, int wave) {
    double run[NTAB], inc[NTAB];
#pragma unroll
    for (int t = 0; t < NTAB; t++) {
        run[t] = 0;
#pragma unroll
        for (int i = 0; i < M; i++)
            if (FULL || i < m) run[t] += val[t][i];
        inc[t] = wave_scan_incl(run[t], lane);
        if (lane == 63) ws[t * 4 + wave] = inc[t];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NTAB; t++) {
        LEnt<T>* tt = tb + t * tabn;
        // the four wave totals, summed over the waves below this one (wave is uniform)
        const double w0 = ws[t * 4], w1 = ws[t * 4 + 1], w2 = ws[t * 4 + 2], w3 = ws[t * 4 + 3];
        const double base = (wave > 0 ? w0 : 0.0) + (wave > 1 ? w1 : 0.0) + (wave > 2 ? w2 : 0.0);
        double pref = base + inc[t] - run[t];  // sum of all pixels before p0
#pragma unroll
        for (int i = 0; i < M; i++) {
            if (FULL || i < m) {
                const int p = p0 + i;
                if (FULL || p < n) lent_write(tt + p + 1, pref - (double)(p + 1) * val[t][i], val[t][i]);
                pref += val[t][i];
            }
        }
        if (tid == 0) lent_write(tt, 0.0, 0.0);
        if (tid == NT - 1) lent_write(tt + n + 1, w0 + w1 + w2 + w3, 0.0);
    }
}

// LDS bytes of one fwd_block with G views per block.
__host__ __device__ inline size_t fwd_block_lds(int nmax, int G, bool weighted, int entry_bytes) {
    const int ntab = weighted ? G : 1;
    return (size_t)2 * ntab * (nmax + 2) * entry_bytes + (size_t)2 * G * LC_WORDS * 8 + (size_t)2 * ntab * 4 * 8;
}

// ---- Line tables built once per (slice, line), then streamed into LDS ----
// Per line TABN = 256*M + 2 entries {C_j, S_j} (16 bytes, j = 0..n+1 used), the same
// layout in global memory and in LDS, so one linear copy moves a table.
template <int M>
__host__ __device__ constexpr int ltab_n() { return NT * M + 2; }

// One line's table by one wave (n <= 256*M): in round k lane L takes pixel p = 64 k + L, a wave
// scan gives the pixels before it, and the 64 lanes store 64 consecutive entries (1 KB, fully
// coalesced); the entries are those line_tables writes for the same line.
template <typename T, int M>
__device__ __forceinline__ void line_table_wave(const T* __restrict__ row, int n, double2* __restrict__ out, int lane) {
    double carry = 0.0;  // sum of the pixels of the earlier rounds
#pragma unroll
    for (int k = 0; k < 4 * M; k++) {
        if (64 * k >= n) break;  // wave-uniform
        const int p = 64 * k + lane;
        const double v = p < n ? (double)row[p] : 0.0;
        const double inc = wave_scan_incl(v, lane);
        if (p < n) out[p + 1] = make_double2(carry + (inc - v) - (double)(p + 1) * v, v);
        carry += readlane_d(inc, 63);
    }
    if (lane == 0) {
        out[0] = make_double2(0.0, 0.0);
        out[n + 1] = make_double2(carry, 0.0);
    }
}

// One line -> bin work item: batch element b, sweep axis `mode`, views
// vlist[grp*G .. grp*G+G-1] of that axis (G = 4*VPW).  `src` holds the lines
// of that axis.  Wave w owns views w*VPW .. w*VPW+VPW-1 and, for each, the bin
// edges k = lane + 64*i (i < NEM).  Per image line: the line is loaded one
// iteration ahead into registers, its cumulative table is built in LDS
// (double-buffered) by the whole block, and the G views' LineCoefs for the line
// (prefetched one line ahead by G*8 threads with vector loads) are staged in
// LDS next to it; each wave then reads only its own views' coefficients.
// The bin values y_k = post_k (S_{k+1} - S_k) are formed with wave shuffles.
// RESID: also write the residual bin tables of projection_t (g = (y - m) post).
// FULL: every line of both axes has exactly 256*M pixels (no pixel predicates).
// EDGEX (unweighted): D == 64*NEM, so the last bin edge D is the only edge of
// group NEM; lane v of the wave evaluates it for view v (VPW <= 64), one
// evaluation per line for all the wave's views instead of a mostly idle group.
// Only lines [lb, le) of the axis contribute (le < 0: all): a launch split into
// line chunks writes each chunk's partial sinogram and sums them afterwards.
template <typename T, int VPW, int NEM, int M, bool WEIGHTED, bool RESID, bool FULL = false, bool EDGEX = false>
__device__ __forceinline__ void fwd_block(const DevPlan& P, const T* __restrict__ src, T* __restrict__ y, int b,
                                          int mode, int grp, const double* __restrict__ post,
                                          const T* __restrict__ m, double2* __restrict__ tabs,
                                          unsigned char* __restrict__ smem, int lb = 0, int le = -1,
                                          double* __restrict__ rowsq = nullptr) {
    constexpr int G = 4 * VPW;
    constexpr int NTAB = WEIGHTED ? G : 1;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int* vlist = mode == MODE_Y ? P.yviews : P.xviews;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const int nlines = mode == MODE_Y ? P.W : P.H;
    const int lend = le < 0 ? nlines : le;
    const int n = FULL ? NT * M : (mode == MODE_Y ? P.H : P.W);
    const int D = P.D;

    int vw[VPW];
#pragma unroll
    for (int v = 0; v < VPW; v++) {
        const int idx = grp * G + wave * VPW + v;
        vw[v] = vlist[idx < nv ? idx : nv - 1];
    }
    double e[VPW][NEM], S[VPW][NEM];
#pragma unroll
    for (int v = 0; v < VPW; v++) {
#pragma unroll
        for (int i = 0; i < NEM; i++) {
            const int k = lane + 64 * i;
            e[v][i] = P.edges[(size_t)vw[v] * (D + 1) + (k <= D ? k : D)];
            S[v][i] = 0.0;
        }
    }
    // EDGEX: lane v (< VPW) owns edge D of view v; the other lanes repeat view 0's
    const int vxl = lane < VPW ? lane : 0;
    int vx = vw[0];
#pragma unroll
    for (int v = 1; v < VPW; v++)
        if (vxl == v) vx = vw[v];
    const double ex = EDGEX ? P.edges[(size_t)vx * (D + 1) + D] : 0.0;
    double Sx = 0.0;

    const int tabn = n + 2;
    LEnt<T>* tab = reinterpret_cast<LEnt<T>*>(smem);                // [2][NTAB][tabn]
    double* cf = reinterpret_cast<double*>(tab + 2 * NTAB * tabn);  // [2][G][LC_WORDS]
    double* wsum = cf + 2 * G * LC_WORDS;                           // [2][NTAB][4]
    const T* base = src + (size_t)b * nlines * n;
    const int m_px = FULL ? M : (n + NT - 1) / NT;
    const int p0 = tid * m_px;
    const int hi = n + 1;

    // coefficient word `cw` of view `cg` for the next line, held by thread tid < G*8
    const bool cthread = tid < G * LC_WORDS;
    const int cg = tid / LC_WORDS, cw = tid % LC_WORDS;
    const double* crow = nullptr;
    if (cthread) {
        const int idx = grp * G + cg;
        crow = reinterpret_cast<const double*>(P.lines + (size_t)vlist[idx < nv ? idx : nv - 1] * P.max_lines) + cw;
    }
    const bool any = lb < lend;
    double cnext = cthread && any ? crow[(size_t)lb * LC_WORDS] : 0.0;

    T xv[M];
    const T* lb0 = base + (size_t)lb * n;
#pragma unroll
    for (int i = 0; i < M; i++) xv[i] = any && (FULL || (i < m_px && p0 + i < n)) ? lb0[p0 + i] : T(0);

    for (int l = lb; l < lend; l++) {
        LEnt<T>* tb = tab + (l & 1) * NTAB * tabn;
        double* ws = wsum + (l & 1) * NTAB * 4;
        double* cfl = cf + (l & 1) * G * LC_WORDS;
        if (cthread) cfl[tid] = cnext;
        if (WEIGHTED) __syncthreads();  // the weights below read cfl
        double val[NTAB][M];
#pragma unroll
        for (int t = 0; t < NTAB; t++) {
            if (WEIGHTED) {
                const double w0 = cfl[t * LC_WORDS + LC_W0], w1 = cfl[t * LC_WORDS + LC_W1];
                const double w2 = cfl[t * LC_WORDS + LC_W2], is = cfl[t * LC_WORDS + LC_INVS];
#pragma unroll
                for (int i = 0; i < M; i++) {
                    val[t][i] = 0.0;
                    if (FULL || (i < m_px && p0 + i < n)) {
                        const double u = __builtin_fma(w1, (double)(p0 + i), w0);
                        const double d = __builtin_fma(u, u, w2);
                        // rounded to the slope's storage precision before the prefix sum
                        const double rd = sizeof(T) == 8 ? rcp_refined(d) : (double)__builtin_amdgcn_rcpf((float)d);
                        val[t][i] = (double)(T)((double)xv[i] * (is * rd));
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < M; i++) val[t][i] = (FULL || (i < m_px && p0 + i < n)) ? (double)xv[i] : 0.0;
            }
        }
        line_tables<T, M, NTAB, FULL>(val, m_px, p0, n, tb, tabn, ws, tid, lane, wave);
        if (l + 1 < lend) {
            const T* nb = base + (size_t)(l + 1) * n;
#pragma unroll
            for (int i = 0; i < M; i++) xv[i] = (FULL || (i < m_px && p0 + i < n)) ? nb[p0 + i] : T(0);
            if (cthread) cnext = crow[(size_t)(l + 1) * LC_WORDS];
        }
        __syncthreads();
        if (EDGEX) {
            const int g = wave * VPW + vxl;  // per lane
            const double tau = __builtin_fma(cfl[g * LC_WORDS + LC_A], ex, cfl[g * LC_WORDS + LC_B]);
            double c, sl;
            lent_read(tb + table_index_d(tau, hi), c, sl);
            Sx = __builtin_fma(cfl[g * LC_WORDS + LC_S], __builtin_fma(tau, sl, c), Sx);
        }
#pragma unroll
        for (int v = 0; v < VPW; v++) {
            const int g = wave * VPW + v;
            const double s = cfl[g * LC_WORDS + LC_S];
            const double A = cfl[g * LC_WORDS + LC_A];
            const double Bc = cfl[g * LC_WORDS + LC_B];
            const LEnt<T>* t = tb + (WEIGHTED ? g : 0) * tabn;
            // every slot is evaluated (slots past the last edge repeat edge D): no branch
            // between the independent evaluations, so their LDS reads overlap
#pragma unroll
            for (int i = 0; i < NEM; i++) {
                const double tau = __builtin_fma(A, e[v][i], Bc);
                double c, sl;
                lent_read(t + table_index_d(tau, hi), c, sl);
                S[v][i] = __builtin_fma(s, __builtin_fma(tau, sl, c), S[v][i]);
            }
        }
    }

    fwd_store<T, VPW, NEM, RESID, EDGEX>(P, S, vw, grp, b, mode, lane, wave, y, post, m, tabs, Sx, rowsq);
}

// LDS bytes of one fwd_block_lt: NB line-table buffers and NB LineCoef buffers.
template <int M>
__host__ __device__ inline size_t fwd_lt_lds(int G, int NB) {
    return (size_t)NB * ltab_n<M>() * 16 + (size_t)NB * G * LC_WORDS * 8;
}

template <int N>
__device__ __forceinline__ void wait_vm_lgkm() {
    asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)" ::"n"(N) : "memory");
}

// fwd_block over prebuilt line tables (unweighted; k_line_tables): `lt` holds the axis's
// lines, ltab_n<M>() entries each, so a line costs no scan and one barrier, and `src` is not
// needed.  Line l's table and the G LineCoefs of line l are copied global -> LDS by
// global_load_lds_dwordx4 (16 bytes per lane straight into LDS, no register staging; the
// LineCoefs by wave 0's lanes < 4G) NB - 1 lines ahead into buffer l % NB.  NB = 2 is used:
// three buffers (two lines ahead) measured 2-4 % slower on 1024-pixel lines, where their LDS
// lowers the blocks per CU.  Per line one wait and one barrier: each wave waits until at most
// the copies of the lines after l are outstanding (loads complete in order), then the barrier
// makes every wave's copies of line l visible and frees the buffer of line l - 1.  Wave w
// issues ceil((TABN - 64 w) / 256) table copies per line (M + 1 for wave 0, M for the
// others) and wave 0 one LineCoef copy; the waits count exactly those.
template <typename T, int VPW, int NEM, int M, bool RESID, bool EDGEX, int NB>
__device__ __forceinline__ void fwd_block_lt(const DevPlan& P, const double2* __restrict__ lt, T* __restrict__ y,
                                             int b, int mode, int grp, const double* __restrict__ post,
                                             const T* __restrict__ m, double2* __restrict__ tabs,
                                             unsigned char* __restrict__ smem, int lb, int le,
                                             double* __restrict__ rowsq) {
    static_assert(NB == 2 || NB == 3, "two or three buffers");
    constexpr int G = 4 * VPW;
    constexpr int TABN = ltab_n<M>();
    static_assert(4 * G <= 64, "one LineCoef copy instruction per line");
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int* vlist = mode == MODE_Y ? P.yviews : P.xviews;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const int nlines = mode == MODE_Y ? P.W : P.H;
    const int lend = le < 0 ? nlines : le;
    const int n = mode == MODE_Y ? P.H : P.W;
    const int D = P.D;

    int vw[VPW];
#pragma unroll
    for (int v = 0; v < VPW; v++) {
        const int idx = grp * G + wave * VPW + v;
        vw[v] = vlist[idx < nv ? idx : nv - 1];
    }
    double e[VPW][NEM], S[VPW][NEM];
#pragma unroll
    for (int v = 0; v < VPW; v++) {
#pragma unroll
        for (int i = 0; i < NEM; i++) {
            const int k = lane + 64 * i;
            e[v][i] = P.edges[(size_t)vw[v] * (D + 1) + (k <= D ? k : D)];
            S[v][i] = 0.0;
        }
    }
    const int vxl = lane < VPW ? lane : 0;
    int vx = vw[0];
#pragma unroll
    for (int v = 1; v < VPW; v++)
        if (vxl == v) vx = vw[v];
    const double ex = EDGEX ? P.edges[(size_t)vx * (D + 1) + D] : 0.0;
    double Sx = 0.0;

    LEnt<T>* tab = reinterpret_cast<LEnt<T>*>(smem);             // [NB][TABN]
    double* cf = reinterpret_cast<double*>(tab + NB * TABN);    // [NB][G][LC_WORDS]
    const int hi = n + 1;
    const bool clane = wave == 0 && lane < 4 * G;
    const char* csrc = nullptr;
    if (clane) {
        const int idx = grp * G + (lane >> 2);
        csrc = reinterpret_cast<const char*>(P.lines + (size_t)vlist[idx < nv ? idx : nv - 1] * P.max_lines) +
               16 * (lane & 3);
    }
    auto issue = [&](int l, int buf) {
        const double2* srct = lt + (size_t)l * TABN;
        LEnt<T>* dst = tab + buf * TABN;
#pragma unroll
        for (int r = 0; r < (TABN + NT - 1) / NT; r++) {
            const int j0 = wave * 64 + r * NT;
            if (j0 < TABN && j0 + lane < TABN)
                __builtin_amdgcn_global_load_lds((const void*)(srct + j0 + lane),
                                                 (__attribute__((address_space(3))) void*)(dst + j0), 16, 0, 0);
        }
        if (clane)
            __builtin_amdgcn_global_load_lds((const void*)(csrc + (size_t)l * sizeof(LineCoef)),
                                             (__attribute__((address_space(3))) void*)(cf + buf * G * LC_WORDS), 16,
                                             0, 0);
    };
#pragma unroll
    for (int k = 0; k < NB - 1; k++)
        if (lb + k < lend) issue(lb + k, k);

    int bcur = 0;
    for (int l = lb; l < lend; l++) {
        // copies of line l complete (at most those of line l + 1 outstanding when NB = 3)
        if (NB == 3 && l + 1 < lend) {
            if (wave == 0)
                wait_vm_lgkm<M + 2>();
            else
                wait_vm_lgkm<M>();
        } else {
            wait_vm_lgkm<0>();
        }
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (l + NB - 1 < lend) issue(l + NB - 1, bcur == 0 ? NB - 1 : bcur - 1);
        const LEnt<T>* tb = tab + bcur * TABN;
        const double* cfl = cf + bcur * G * LC_WORDS;
        if (EDGEX) {
            const int g = wave * VPW + vxl;
            const double tau = __builtin_fma(cfl[g * LC_WORDS + LC_A], ex, cfl[g * LC_WORDS + LC_B]);
            double c, sl;
            lent_read(tb + table_index_d(tau, hi), c, sl);
            Sx = __builtin_fma(cfl[g * LC_WORDS + LC_S], __builtin_fma(tau, sl, c), Sx);
        }
#pragma unroll
        for (int v = 0; v < VPW; v++) {
            const int g = wave * VPW + v;
            const double s = cfl[g * LC_WORDS + LC_S];
            const double A = cfl[g * LC_WORDS + LC_A];
            const double Bc = cfl[g * LC_WORDS + LC_B];
#pragma unroll
            for (int i = 0; i < NEM; i++) {
                const double tau = __builtin_fma(A, e[v][i], Bc);
                double c, sl;
                lent_read(tb + table_index_d(tau, hi), c, sl);
                S[v][i] = __builtin_fma(s, __builtin_fma(tau, sl, c), S[v][i]);
            }
        }
        bcur = bcur == NB - 1 ? 0 : bcur + 1;
    }

    fwd_store<T, VPW, NEM, RESID, EDGEX>(P, S, vw, grp, b, mode, lane, wave, y, post, m, tabs, Sx, rowsq);
}

// Entries of a wave-private line table sit at j + j/8 (one 16-byte pad every 8):
// lane L writes entries 8L+1 .. 8L+8, and with the pad consecutive lanes land
// 9 entries apart, so each ds_write_b128 of 16 lanes covers all 64 banks.
__host__ __device__ inline int wtab_slots(int n) { return n + 2 + (n + 2) / 8 + 1; }
__device__ __forceinline__ int wtab_slot(int j) { return j + (j >> 3); }

// LDS bytes of one fwd_block_w: one wave-private table per wave.
__host__ __device__ inline size_t fwd_w_block_lds(int nmax) { return (size_t)4 * wtab_slots(nmax) * 16; }

// Weighted line -> bin work item (backprojection_t and its exact adjoint form)
// for lines of at most 64*PPL pixels: G = 4 views per block, one per wave, and
// each wave builds the weighted cumulative table of its own view for every line
// by itself: lane L owns pixels L*PPL .. L*PPL+PPL-1 (one serial sum and one
// wave scan per line) and writes them into its wave's private table, so no
// barrier is needed anywhere.  The line's pixels and the view's LineCoef words
// are loaded one line ahead into registers.  FULL: every line has exactly
// 64*PPL pixels, so no lane needs a bounds predicate.  The pixel weight
// 1/(s ((w0 + w1 p)^2 + w2)) is formed in the data precision (fp32 data: fp32
// arithmetic with a fp32 reciprocal, about 1e-7 relative; fp64 data: fp64
// arithmetic with a Newton-refined reciprocal), and the weighted value is rounded to the
// data type before the fp64 prefix sum, as in fwd_block with WEIGHTED.
// Lines [lb, le) of the axis contribute (le < 0: all), as in fwd_block.
template <typename T, int NEM, int PPL, bool FULL>
__device__ __forceinline__ void fwd_block_w(const DevPlan& P, const T* __restrict__ src, T* __restrict__ y, int b,
                                            int mode, int grp, const double* __restrict__ post,
                                            unsigned char* __restrict__ smem, int lb = 0, int le = -1) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int* vlist = mode == MODE_Y ? P.yviews : P.xviews;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const int nlines = mode == MODE_Y ? P.W : P.H;
    const int n = FULL ? 64 * PPL : (mode == MODE_Y ? P.H : P.W);
    const int D = P.D;

    int vw[1];
    {
        const int idx = grp * 4 + wave;
        vw[0] = vlist[idx < nv ? idx : nv - 1];
    }
    double e[1][NEM], S[1][NEM];
#pragma unroll
    for (int i = 0; i < NEM; i++) {
        const int k = lane + 64 * i;
        e[0][i] = P.edges[(size_t)vw[0] * (D + 1) + (k <= D ? k : D)];
        S[0][i] = 0.0;
    }

    double2* tb = reinterpret_cast<double2*>(smem) + (size_t)wave * wtab_slots(n);
    const T* base = src + (size_t)b * nlines * n;
    const int p0 = lane * PPL;
    const int hi = n + 1;
    // LineCoef word `lane` (lanes 0..7) of this wave's view, one line ahead
    const double* crow = reinterpret_cast<const double*>(P.lines + (size_t)vw[0] * P.max_lines) + (lane & 7);
    // fp32 data: the line's pre-scaled weight coefficients {., c w1, c w0, c^2 w2} (uniform)
    const float4* wrow = P.adj_wk + (size_t)vw[0] * P.max_lines;
    auto load_line = [&](int l, T (&xv)[PPL], double& cw, float4& wk) {
        const T* row = base + (size_t)l * n;
#pragma unroll
        for (int k = 0; k < PPL; k++) xv[k] = (FULL || p0 + k < n) ? row[p0 + k] : T(0);
        cw = crow[(size_t)l * LC_WORDS];
        if (sizeof(T) == 4) wk = wrow[l];
    };
    const int lend = le < 0 ? nlines : le;
    T xv[PPL];
    double cw = 0.0;
    float4 wk = make_float4(0.f, 0.f, 0.f, 1.f);
#pragma unroll
    for (int k = 0; k < PPL; k++) xv[k] = T(0);
    if (lb < lend) load_line(lb, xv, cw, wk);

    for (int l = lb; l < lend; l++) {
        const double s = readlane_d(cw, LC_S), A = readlane_d(cw, LC_A), Bc = readlane_d(cw, LC_B);
        const double w0 = readlane_d(cw, LC_W0), w1 = readlane_d(cw, LC_W1);
        const double w2 = readlane_d(cw, LC_W2), is = readlane_d(cw, LC_INVS);
        double val[PPL];
        if constexpr (sizeof(T) == 4) {
            // weight (1/s) / ((w0 + w1 p)^2 + w2) = 1 / (u^2 + c^2 w2), u = k c w1 + (p0 c w1 + c w0)
            const float u0 = __builtin_fmaf(wk.y, (float)p0, wk.z);
#pragma unroll
            for (int k = 0; k < PPL; k++) {
                const float u = k == 0 ? u0 : __builtin_fmaf((float)k, wk.y, u0);
                const float v = (float)xv[k] * __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, wk.w));
                val[k] = (FULL || p0 + k < n) ? (double)v : 0.0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < PPL; k++) {
                const double u = __builtin_fma(w1, (double)(p0 + k), w0);
                const double d = __builtin_fma(u, u, w2);
                const double v = (double)xv[k] * (is * rcp_refined(d));
                val[k] = (FULL || p0 + k < n) ? v : 0.0;
            }
        }
        double run = 0.0;
#pragma unroll
        for (int k = 0; k < PPL; k++) run += val[k];
        if (l + 1 < lend) load_line(l + 1, xv, cw, wk);
        // Table of this line: the wave's reads of the previous line's table were
        // consumed above (in-order LDS within a wave), so it can be overwritten.
        const double inc = wave_scan_incl(run, lane);
        double pref = inc - run;
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            const int p = p0 + k;
            if (FULL || p < n) tb[wtab_slot(p + 1)] = make_double2(__builtin_fma(-(double)(p + 1), val[k], pref), val[k]);
            pref += val[k];
        }
        if (lane == 0) tb[0] = make_double2(0.0, 0.0);
        if (lane == 63) tb[wtab_slot(n + 1)] = make_double2(inc, 0.0);
#pragma unroll
        for (int i = 0; i < NEM; i++) {
            const double tau = __builtin_fma(A, e[0][i], Bc);
            const double2 ent = tb[wtab_slot(table_index_d(tau, hi))];
            S[0][i] = __builtin_fma(s, __builtin_fma(tau, ent.y, ent.x), S[0][i]);
        }
    }
    fwd_store<T, 1, NEM, false>(P, S, vw, grp, b, mode, lane, wave, y, post, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// Bin -> pixel (projection_t / backprojection)
// ---------------------------------------------------------------------------

// Axis coordinate -> bin lookup of one view (uniform values, held in SGPRs).
struct Lookup {
    float lk[4];      // Moebius (ED, EA) or affine (parallel) coefficients
    float ea_sc, ea_off;
    float c[6], T;    // fan-EA odd polynomial for atan on |t| <= T (GEO_EAP)
    double lkd[6];    // lk0..lk3, ea_sc, ea_off in fp64 (fp64 data)
};

__device__ __forceinline__ Lookup view_lookup(const DevPlan& P, int v) {
    const ViewParams& vp = P.vp[v];
    Lookup L;
#pragma unroll
    for (int k = 0; k < 4; k++) L.lk[k] = vp.lk[k];
    L.ea_sc = vp.ea_sc;
    L.ea_off = vp.ea_off;
#pragma unroll
    for (int k = 0; k < 6; k++) L.c[k] = P.ea_c[k];
    L.T = P.ea_T;
    const double* d = P.lookup_d + (size_t)v * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) L.lkd[k] = d[k];
    return L;
}

// G(pi_p) of pixel edge p of a line with edges pi_p = a + s*p.  GEO_EA takes the
// arctangent with atanf; GEO_EAP with the plan's odd polynomial (t clamped to
// [-T, T], beyond which the index clamps to an end entry anyway).  F64Q (fp64
// data): the table index comes from the fp64 coefficients in fp64 arithmetic
// (fan-EA with atan), since an fp32 index near a bin boundary picks the
// neighbouring segment, an error of ~1e-4 bin widths times the jump in slope.
template <int GEO, bool F64Q = false>
__device__ __forceinline__ double eval_G(const double2* __restrict__ t, double a, double s, int p, const Lookup& L,
                                         float hi) {
    const double pi = __builtin_fma(s, (double)p, a);
    if (F64Q) {
        double qd;
        if (GEO == GEO_PARA) {
            qd = __builtin_fma(L.lkd[0], pi, L.lkd[1]);
        } else {
            const double num = __builtin_fma(L.lkd[0], pi, L.lkd[1]);
            const double den = __builtin_fma(L.lkd[2], pi, L.lkd[3]);
            qd = num * rcp_refined(den);
            if (GEO == GEO_EA || GEO == GEO_EAP) qd = __builtin_fma(atan(qd), L.lkd[4], L.lkd[5]);
        }
        const double2 c = t[table_index_d(qd, (int)hi)];
        return __builtin_fma(pi, c.y, c.x);
    }
    const float pf = (float)pi;
    float q;
    if (GEO == GEO_PARA) {
        q = __builtin_fmaf(L.lk[0], pf, L.lk[1]);
    } else {
        const float num = __builtin_fmaf(L.lk[0], pf, L.lk[1]);
        const float den = __builtin_fmaf(L.lk[2], pf, L.lk[3]);
        q = num * __builtin_amdgcn_rcpf(den);
        if (GEO == GEO_EA) q = __builtin_fmaf(atanf(q), L.ea_sc, L.ea_off);
        if (GEO == GEO_EAP) {
            const float tc = __builtin_amdgcn_fmed3f(q, -L.T, L.T);
            const float s2 = tc * tc;
            float z = L.c[5];
#pragma unroll
            for (int k = 4; k >= 0; k--) z = __builtin_fmaf(z, s2, L.c[k]);
            q = __builtin_fmaf(tc * z, L.ea_sc, L.ea_off);
        }
    }
    const double2 c = t[table_index(q, hi)];
    return __builtin_fma(pi, c.y, c.x);
}

// The lookup of one view composed with one line's pixel-edge map, fp32 data.  The
// plan stores {64 c0, c1, 64 c2, c3} (HostPlan::adj_lk holds c0..c3 scaled so):
// at pixel edge p = lane + 64*i the lookup argument is (n0 + i C0) / (d0 + i C2)
// with n0 = C0 lane/64 + c1, d0 = C2 lane/64 + c3 (one fp32 FMA each with the
// literal i), and the axis coordinate pi0 + i (64 s) is stepped by one fp64 add
// per slot.  No edge position is converted to fp32 and no per-slot pixel index
// is held in registers.
struct LineLk {
    double pi0, s64;
    float n0, C0, d0, C2;
};
__device__ __forceinline__ LineLk line_lk(double a, double s, float4 c, int lane) {
    LineLk k;
    k.pi0 = __builtin_fma(s, (double)lane, a);
    k.s64 = 64.0 * s;
    const float lq = (float)lane * (1.0f / 64.0f);  // exact
    k.n0 = __builtin_fmaf(c.x, lq, c.y);
    k.C0 = c.x;
    k.d0 = __builtin_fmaf(c.z, lq, c.w);
    k.C2 = c.z;
    return k;
}

// eval_G of pixel edge lane + 64*i of a line at axis coordinate pi through its
// LineLk (fp32 data; the bin lookup of eval_G's fp32 path from the composed
// coefficients).
template <int GEO>
__device__ __forceinline__ double eval_G_lk(const double2* __restrict__ t, double pi, const LineLk& k, int i,
                                            const Lookup& L, float hi) {
    const float num = i == 0 ? k.n0 : __builtin_fmaf((float)i, k.C0, k.n0);
    float q;
    if (GEO == GEO_PARA) {
        q = num;
    } else {
        const float den = i == 0 ? k.d0 : __builtin_fmaf((float)i, k.C2, k.d0);
        q = num * __builtin_amdgcn_rcpf(den);
        if (GEO == GEO_EA) q = __builtin_fmaf(atanf(q), L.ea_sc, L.ea_off);
        if (GEO == GEO_EAP) {
            const float tc = __builtin_amdgcn_fmed3f(q, -L.T, L.T);
            const float s2 = tc * tc;
            float z = L.c[5];
#pragma unroll
            for (int j = 4; j >= 0; j--) z = __builtin_fmaf(z, s2, L.c[j]);
            q = __builtin_fmaf(tc * z, L.ea_sc, L.ea_off);
        }
    }
    const double2 c = t[table_index(q, hi)];
    return __builtin_fma(pi, c.y, c.x);
}

// LDS bytes of one adj_block.
__host__ __device__ inline size_t adj_block_lds(int D, int L) {
    return (size_t)2 * (D + 2) * 16 + (size_t)2 * L * LC_WORDS * 8 + (size_t)4 * L * 16;
}

// One bin -> pixel work item: batch element b, sweep axis `mode`, image lines
// [tile*L, tile*L + L) (L = 4*LPW; rows for X, columns bottom-to-top for Y)
// over all views of that axis.  X lines are written to outX [B][H][W], Y lines
// to outYT [B][W][H].  Per view: the next view's bin table is copied
// global -> LDS with global_load_lds_dwordx4 (asynchronous, no register
// staging) and the next view's LineCoefs of the tile are prefetched into one
// register per thread while the current view is evaluated; one barrier per view.
// EXACT: every line has exactly 64*(NPE-1) pixels, so its last pixel edge n is
// the only edge of group NPE-1; instead of a ninth (or seventeenth) group per
// line, lane q of the wave evaluates edge n of its line q (LPW <= 64 lanes),
// one evaluation per view for all the wave's lines.
// Views [vb, ve) of the axis list (ve < 0: all) contribute; a launch split into
// view chunks writes each chunk's partial image and sums them afterwards.
template <typename T, int LPW, int NPE, bool WEIGHTED, int GEO, bool EXACT = false>
__device__ __forceinline__ void adj_block(const DevPlan& P, const double2* __restrict__ tabs, T* __restrict__ outX,
                                          T* __restrict__ outYT, int b, int mode, int tile, double wscale,
                                          unsigned char* __restrict__ smem, int vb = 0, int ve = -1) {
    constexpr int L = 4 * LPW;
    constexpr int NG = EXACT ? NPE - 1 : NPE;  // edge groups evaluated per line
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int* vlist = mode == MODE_Y ? P.yviews : P.xviews;
    const int nv = ve < 0 ? (mode == MODE_Y ? P.ny : P.nx) : ve;
    const int nlines = mode == MODE_Y ? P.W : P.H;
    const int n = mode == MODE_Y ? P.H : P.W;
    const int D = P.D;
    const int npe = (n + 1 + 63) / 64;  // edge groups (<= NPE)
    const float hi = (float)(D + 1);
    const int tabn = D + 2;

    double2* stab = reinterpret_cast<double2*>(smem);          // [2][tabn]
    double* scf = reinterpret_cast<double*>(stab + 2 * tabn);  // [2][L][LC_WORDS]
    float4* slk = reinterpret_cast<float4*>(scf + 2 * L * LC_WORDS);  // [2][L] composed lookups (fp32 data)
    float4* swk = slk + 2 * L;  // [2][L] pre-scaled pixel-weight coefficients (fp32 data, weighted)
    constexpr bool LK = sizeof(T) == 4;
    constexpr bool WK = WEIGHTED && LK;

    const double2* tb_b = tabs + (size_t)b * P.V * tabn;
    // Asynchronous copy of view v's table into LDS buffer `buf`: wave w copies
    // entries [64*(w + 4r), 64*(w + 4r) + 64); lanes past the end are masked off.
    auto table_to_lds = [&](int v, int buf) {
        const double2* src = tb_b + (size_t)v * tabn;
        for (int j0 = wave * 64; j0 < tabn; j0 += NT) {
            if (j0 + lane < tabn)
                __builtin_amdgcn_global_load_lds((const void*)(src + j0 + lane),
                                                 (__attribute__((address_space(3))) void*)(stab + buf * tabn + j0), 16,
                                                 0, 0);
        }
    };
    // LineCoef word `tid % 8` of tile line `tid / 8`, held by thread tid < L*8.
    const bool cthread = tid < L * LC_WORDS;
    const int cl = tile * L + tid / LC_WORDS;
    const bool cvalid = cthread && cl < nlines;
    auto coef_word = [&](int v) {
        return cvalid ? reinterpret_cast<const double*>(P.lines + (size_t)v * P.max_lines + cl)[tid % LC_WORDS] : 0.0;
    };
    // composed lookup of tile line `tid`, held by thread tid < L (fp32 data)
    const bool athread = LK && tid < L;
    const int al = tile * L + tid;
    auto lk_word = [&](int v) {
        return athread && al < nlines ? P.adj_lk[(size_t)v * P.max_lines + al] : make_float4(0.f, 0.f, 0.f, 1.f);
    };
    auto wk_word = [&](int v) {
        return athread && al < nlines ? P.adj_wk[(size_t)v * P.max_lines + al] : make_float4(0.f, 0.f, 0.f, 1.f);
    };

    // Tiles of 16 or more lines (DMA): the LineCoefs (16 bytes of line tid/4 per thread
    // tid < 4L) and composed lookups (thread tid < L) are copied global -> LDS like the table,
    // and each view ends with the wait for this wave's copies and a bare barrier instead of
    // register-staged words and __syncthreads() (whose workgroup fence drains every access):
    // projection_t 3-4 % faster at fan-ED 512^2, 1-2 % on fan-EA; 8-line tiles (lines over 512
    // pixels) measured 1.5-2.5 % slower that way and keep the staged words.
    constexpr bool DMA = LPW >= 4;
    auto side_to_lds = [&](int v, int buf) {
        if (tid < 4 * L) {
            const int c4 = min(tile * L + (tid >> 2), nlines - 1);
            __builtin_amdgcn_global_load_lds(
                (const void*)(reinterpret_cast<const char*>(P.lines + (size_t)v * P.max_lines + c4) + 16 * (tid & 3)),
                (__attribute__((address_space(3))) void*)(scf + buf * L * LC_WORDS + wave * 128), 16, 0, 0);
        }
        if (LK && tid < L)
            __builtin_amdgcn_global_load_lds((const void*)(P.adj_lk + (size_t)v * P.max_lines + min(al, nlines - 1)),
                                             (__attribute__((address_space(3))) void*)(slk + buf * L), 16, 0, 0);
        if (WK && tid < L)
            __builtin_amdgcn_global_load_lds((const void*)(P.adj_wk + (size_t)v * P.max_lines + min(al, nlines - 1)),
                                             (__attribute__((address_space(3))) void*)(swk + buf * L), 16, 0, 0);
    };
    auto view_barrier = [&]() {
        if (DMA) {
            asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
        } else {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
    };

    double acc[LPW][NG];
#pragma unroll
    for (int q = 0; q < LPW; q++)
#pragma unroll
        for (int i = 0; i < NG; i++) acc[q][i] = 0.0;
    double accx = 0.0;                          // EXACT: lane q accumulates edge n of line q
    const int lx = (lane < LPW ? lane : 0) * 4 + wave;  // that line within the tile

    table_to_lds(vlist[vb], 0);
    if (DMA) {
        side_to_lds(vlist[vb], 0);
    } else {
        if (cthread) scf[tid] = coef_word(vlist[vb]);
        if (athread) slk[tid] = lk_word(vlist[vb]);
        if (WK && athread) swk[tid] = wk_word(vlist[vb]);
    }
    view_barrier();

    for (int iv = vb; iv < nv; iv++) {
        const int cur = (iv - vb) & 1;
        const int v = vlist[iv];
        double cnext = 0.0;
        float4 anext = make_float4(0.f, 0.f, 0.f, 1.f), wnext = anext;
        if (iv + 1 < nv) {
            const int vn = vlist[iv + 1];
            table_to_lds(vn, cur ^ 1);
            if (DMA) {
                side_to_lds(vn, cur ^ 1);
            } else {
                cnext = coef_word(vn);
                if (athread) anext = lk_word(vn);
                if (WK && athread) wnext = wk_word(vn);
            }
        }
        const double2* t = stab + cur * tabn;
        const double* cf = scf + cur * L * LC_WORDS;
        const float4* clk = slk + cur * L;
        const float4* cwk = swk + cur * L;
        const Lookup lu = view_lookup(P, v);
        double gx = 0.0;  // EXACT: G at edge n of line lx (lanes >= LPW: unused)
        if (EXACT) {
            gx = eval_G<GEO, sizeof(T) == 8>(t, cf[lx * LC_WORDS + LC_A0], cf[lx * LC_WORDS + LC_S], n, lu, hi);
            if (!WEIGHTED) accx += gx;
        }
#pragma unroll
        for (int q = 0; q < LPW; q++) {
            const int lq = q * 4 + wave;  // line within the tile
            const int l = tile * L + lq;
            if (l >= nlines) continue;    // wave-uniform
            const double a = cf[lq * LC_WORDS + LC_A0], s = cf[lq * LC_WORDS + LC_S];
            // Every slot is evaluated, also past the last pixel edge (clamped table
            // index, never stored): no branch between the independent evaluations,
            // so their LDS reads overlap.
            // G at pixel edge lane + 64 i of this line, called for i = 0, 1, 2, ... in order
            LineLk lk{};
            if (LK) lk = line_lk(a, s, clk[lq], lane);
            double pir = lk.pi0;  // axis coordinate of edge lane + 64 i (fp32 data)
            auto G_at = [&](int i) {
                if constexpr (LK) {
                    if (i > 0) pir += lk.s64;
                    return eval_G_lk<GEO>(t, pir, lk, i, lu, hi);
                } else {
                    return eval_G<GEO, true>(t, a, s, lane + 64 * i, lu, hi);
                }
            };
            if (!WEIGHTED) {
#pragma unroll
                for (int i = 0; i < NG; i++) acc[q][i] += G_at(i);
            } else {
                // Pixel weight inv_s / ((w0 + w1 p)^2 + w2) of pixel p = lane + 64 k (the call's
                // scale multiplies the output).  fp32 storage: from the plan's pre-scaled fp32
                // coefficients, 1 / (u^2 + c^2 w2) with u = k (64 c w1) + u0 (relative error
                // ~3e-7, below the output's rounding); fp64 storage: u and d in fp64,
                // Newton-refined reciprocal.
                const double w0 = cf[lq * LC_WORDS + LC_W0], w1 = cf[lq * LC_WORDS + LC_W1];
                const double w2 = cf[lq * LC_WORDS + LC_W2];
                const double ws = cf[lq * LC_WORDS + LC_INVS];
                float4 wk = make_float4(0.f, 0.f, 0.f, 1.f);
                if (WK) wk = cwk[lq];
                const float u0 = __builtin_fmaf(wk.y, (float)lane, wk.z);
                auto wgt = [&](int k) {
                    if constexpr (WK) {
                        const float u = k == 0 ? u0 : __builtin_fmaf((float)k, wk.x, u0);
                        return (double)__builtin_amdgcn_rcpf(__builtin_fmaf(u, u, wk.w));
                    } else {
                        const double u = __builtin_fma(w1, (double)(lane + 64 * k), w0);
                        const double d = __builtin_fma(u, u, w2);
                        return ws * rcp_refined(d);
                    }
                };
                // right edge of pixel lane + 64 (i-1): edge lane+1 of group i-1, or for
                // lane 63 edge 0 of group i; both come from the rotations of the groups
                double gprev = 0.0, rprev = 0.0;
#pragma unroll
                for (int i = 0; i < NG; i++) {
                    const double gi = G_at(i);
                    const double ri = rot_next_d(gi);
                    if (i > 0) {
                        const double right = lane == 63 ? ri : rprev;
                        // pixels past the line accumulate unused values (never stored)
                        acc[q][i - 1] += wgt(i - 1) * (right - gprev);
                    }
                    gprev = gi;
                    rprev = ri;
                }
                if (EXACT) {
                    // last group: lane 63's pixel n-1 ends at edge n, evaluated by lane q
                    const double gxq = readlane_d(gx, q);
                    acc[q][NG - 1] += wgt(NG - 1) * ((lane == 63 ? gxq : rprev) - gprev);
                } else {
                    // pixels of the last slot group (lane 63's pixel is past the line since 64 * NPE > n)
                    acc[q][NPE - 1] += wgt(NPE - 1) * (rprev - gprev);
                }
            }
        }
        if (!DMA && iv + 1 < nv && cthread) scf[(cur ^ 1) * L * LC_WORDS + tid] = cnext;
        if (!DMA && iv + 1 < nv && athread) slk[(cur ^ 1) * L + tid] = anext;
        if (!DMA && WK && iv + 1 < nv && athread) swk[(cur ^ 1) * L + tid] = wnext;
        view_barrier();
    }

    // Write pixel values.
#pragma unroll
    for (int q = 0; q < LPW; q++) {
        const int l = tile * L + q * 4 + wave;
        if (l >= nlines) continue;
        T* dst = mode == MODE_Y ? outYT + ((size_t)b * P.W + l) * P.H : outX + ((size_t)b * P.H + l) * P.W;
#pragma unroll
        for (int i = 0; i < NG; i++) {
            if (i >= npe) continue;  // wave-uniform
            const int p = lane + 64 * i;
            double val;
            if (!WEIGHTED) {
                double right = shfl_next_d(acc[q][i]);
                double r0 = 0.0;
#pragma unroll
                for (int k = 0; k < NG; k++)
                    if (k == i + 1) r0 = readlane_d(acc[q][k], 0);
                if (EXACT && i == NG - 1) r0 = readlane_d(accx, q);
                if (lane == 63) right = r0;
                val = right - acc[q][i];
            } else {
                val = acc[q][i] * wscale;
            }
            if (p < n) dst[p] = (T)val;
        }
    }
}

}  // namespace dd
}  // namespace ctlib
