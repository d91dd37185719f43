// Distance-driven projector kernels for gfx950 (MI355X).
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
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace ctlib {

namespace {

constexpr int NT = 256;  // threads per block (4 wave64)

// Blocks are dispatched round-robin over the 8 XCDs (block id % 8).  Give each
// XCD a contiguous range of work items so that blocks sharing an image or a
// sinogram (same batch element) share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks, int nwork) {
    const int per = nblocks >> 3;  // nblocks is a multiple of 8
    const int w = (bid & 7) * per + (bid >> 3);
    return w < nwork ? w : -1;
}

inline int grid_for(int nwork) { return ((nwork + 7) / 8) * 8; }

template <typename T>
__device__ __forceinline__ double ld(const T* p) { return (double)*p; }

// Index of a cumulative table from a float coordinate, clamped to [0, hi].
__device__ __forceinline__ int table_index(float q, float hi) {
    return (int)__builtin_amdgcn_fmed3f(q, 0.0f, hi);
}

__device__ __forceinline__ double shfl_up_d(double v, int off) { return __shfl_up(v, off, 64); }
__device__ __forceinline__ double shfl_down_d(double v, int off) { return __shfl_down(v, off, 64); }
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int2 w = *reinterpret_cast<int2*>(&v);
    int2 r;
    r.x = __builtin_amdgcn_readlane(w.x, lane);
    r.y = __builtin_amdgcn_readlane(w.y, lane);
    return *reinterpret_cast<double*>(&r);
}

// Inclusive wave64 scan.
__device__ __forceinline__ double wave_scan_incl(double v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double t = shfl_up_d(v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Y-mode lines: xT[b][c][i] = x[b][H-1-i][c]
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_to_ylines(const T* __restrict__ x, T* __restrict__ xT, int H, int W) {
    __shared__ T tile[32][33];
    const int b = blockIdx.z;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const T* xb = x + (size_t)b * H * W;
    T* tb = xT + (size_t)b * H * W;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < H && c < W) tile[ty + k][tx] = xb[(size_t)r * W + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int c = c0 + ty + k;           // output line
        const int r = r0 + 31 - tx;          // row read by this lane
        if (c < W && r < H) tb[(size_t)c * H + (H - 1 - r)] = tile[31 - tx][ty + k];
    }
}

// ---------------------------------------------------------------------------
// Line -> bin (projection / backprojection_t)
// ---------------------------------------------------------------------------
constexpr int LC_WORDS = 8;  // doubles per LineCoef: a, s, A, B, w0, w1, w2, inv_s
enum { LC_A0 = 0, LC_S = 1, LC_A = 2, LC_B = 3, LC_W0 = 4, LC_W1 = 5, LC_W2 = 6, LC_INVS = 7 };

// Builds one line table {C_j, slope_j}, j = 0..n+1, from `val` (this thread's
// pixels p0 .. p0+m-1).  `ws` holds 4 wave sums; one barrier inside.
template <int M>
__device__ __forceinline__ void line_table(const double (&val)[M], int m, int p0, int n, double2* __restrict__ tb,
                                           double* __restrict__ ws, int tid, int lane, int wave) {
    double run = 0;
#pragma unroll
    for (int i = 0; i < M; i++)
        if (i < m) run += val[i];
    const double inc = wave_scan_incl(run, lane);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    double base = 0;
    for (int q = 0; q < wave; q++) base += ws[q];
    double pref = base + inc - run;  // sum of all pixels before p0
#pragma unroll
    for (int i = 0; i < M; i++) {
        if (i < m) {
            const int p = p0 + i;
            if (p < n) tb[p + 1] = make_double2(pref - (double)(p + 1) * val[i], val[i]);
            pref += val[i];
        }
    }
    if (tid == 0) tb[0] = make_double2(0.0, 0.0);
    if (tid == NT - 1) tb[n + 1] = make_double2(ws[0] + ws[1] + ws[2] + ws[3], 0.0);
}

// One launch covers both sweep axes: work items [0, B*ngx) are (b, group of G
// X-axis views) and read image rows from `srcx`; the rest are Y-axis groups
// reading the bottom-to-top column copy `srcy`.
//
// Per image line: the line is loaded one iteration ahead into registers, its
// cumulative table is built in LDS (double-buffered), and the G views'
// LineCoefs for the line (prefetched one line ahead by G*8 threads with vector
// loads, so no scalar-cache round trip sits between LDS reads) are staged in
// LDS next to it.  Each thread then evaluates its NE bin edges for every view.
template <typename T, int G, int NE, int M, bool WEIGHTED>
__global__ __launch_bounds__(256) void k_fwd(DevPlan P, const T* __restrict__ srcx, const T* __restrict__ srcy,
                                             T* __restrict__ y, int B, int ngx, int ngy,
                                             const double* __restrict__ post) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NTAB = WEIGHTED ? G : 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nwx = B * ngx;
    int w = xcd_remap(blockIdx.x, gridDim.x, nwx + B * ngy);
    if (w < 0) return;
    const int mode = w < nwx ? MODE_X : MODE_Y;
    if (mode == MODE_Y) w -= nwx;
    const int ngroups = mode == MODE_Y ? ngy : ngx;
    const int b = w / ngroups, grp = w - b * ngroups;
    const T* src = mode == MODE_Y ? srcy : srcx;
    const int* vlist = mode == MODE_Y ? P.yviews : P.xviews;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const int nlines = mode == MODE_Y ? P.W : P.H;
    const int n = mode == MODE_Y ? P.H : P.W;
    const int D = P.D;

    int vg[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int idx = grp * G + g;
        vg[g] = vlist[idx < nv ? idx : nv - 1];
    }
    double e[G][NE], S[G][NE];
#pragma unroll
    for (int g = 0; g < G; g++) {
#pragma unroll
        for (int i = 0; i < NE; i++) {
            const int k = tid + NT * i;
            e[g][i] = P.edges[(size_t)vg[g] * (D + 1) + (k <= D ? k : D)];
            S[g][i] = 0.0;
        }
    }

    const int tabn = n + 2;
    double2* tab = reinterpret_cast<double2*>(smem);               // [2][NTAB][tabn]
    double* cf = reinterpret_cast<double*>(tab + 2 * NTAB * tabn);  // [2][G][LC_WORDS]
    double* wsum = cf + 2 * G * LC_WORDS;                           // [2][NTAB][4]
    const T* base = src + (size_t)b * nlines * n;
    const int m = (n + NT - 1) / NT;
    const int p0 = tid * m;
    const float hi = (float)(n + 1);

    // coefficient word `cw` of view `cg` for the next line, held by thread tid < G*8
    const bool cthread = tid < G * LC_WORDS;
    const int cg = tid / LC_WORDS, cw = tid % LC_WORDS;
    const double* crow = nullptr;
    if (cthread) {
        const int idx = grp * G + cg;
        crow = reinterpret_cast<const double*>(P.lines + (size_t)vlist[idx < nv ? idx : nv - 1] * P.max_lines) + cw;
    }
    double cnext = cthread ? crow[0] : 0.0;

    T xv[M];
#pragma unroll
    for (int i = 0; i < M; i++) xv[i] = (i < m && p0 + i < n) ? base[p0 + i] : T(0);

    for (int l = 0; l < nlines; l++) {
        double2* tb = tab + (l & 1) * NTAB * tabn;
        double* ws = wsum + (l & 1) * NTAB * 4;
        double* cfl = cf + (l & 1) * G * LC_WORDS;
        if (cthread) cfl[tid] = cnext;
        if (WEIGHTED) __syncthreads();  // the weights below read cfl
#pragma unroll
        for (int t = 0; t < NTAB; t++) {
            double val[M];
            if (WEIGHTED) {
                const double w0 = cfl[t * LC_WORDS + LC_W0], w1 = cfl[t * LC_WORDS + LC_W1];
                const double w2 = cfl[t * LC_WORDS + LC_W2], is = cfl[t * LC_WORDS + LC_INVS];
#pragma unroll
                for (int i = 0; i < M; i++) {
                    val[i] = 0.0;
                    if (i < m && p0 + i < n) {
                        const double u = __builtin_fma(w1, (double)(p0 + i), w0);
                        const double d = __builtin_fma(u, u, w2);
                        val[i] = (double)xv[i] * (is * (double)__builtin_amdgcn_rcpf((float)d));
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < M; i++) val[i] = (i < m && p0 + i < n) ? (double)xv[i] : 0.0;
            }
            line_table<M>(val, m, p0, n, tb + t * tabn, ws + t * 4, tid, lane, wave);
        }
        if (l + 1 < nlines) {
            const T* nb = base + (size_t)(l + 1) * n;
#pragma unroll
            for (int i = 0; i < M; i++) xv[i] = (i < m && p0 + i < n) ? nb[p0 + i] : T(0);
            if (cthread) cnext = crow[(size_t)(l + 1) * LC_WORDS];
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; g++) {
            const double s = cfl[g * LC_WORDS + LC_S];
            const double A = cfl[g * LC_WORDS + LC_A];
            const double Bc = cfl[g * LC_WORDS + LC_B];
            const double2* t = tb + (WEIGHTED ? g : 0) * tabn;
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const double tau = __builtin_fma(A, e[g][i], Bc);
                const int j = table_index((float)tau, hi);
                const double2 c = t[j];
                S[g][i] = __builtin_fma(s, __builtin_fma(tau, c.y, c.x), S[g][i]);
            }
        }
    }

    __syncthreads();
    double* sS = reinterpret_cast<double*>(smem);  // [G][D+1]
#pragma unroll
    for (int g = 0; g < G; g++) {
#pragma unroll
        for (int i = 0; i < NE; i++) {
            const int k = tid + NT * i;
            if (k <= D) sS[g * (D + 1) + k] = S[g][i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; g++) {
        if (grp * G + g >= nv) break;
        const int v = vg[g];
        const bool rev = P.vp[v].rev;
        T* yrow = y + ((size_t)b * P.V + v) * D;
        const double* pr = post ? post + (size_t)v * D : nullptr;
        for (int k = tid; k < D; k += NT) {
            double val = sS[g * (D + 1) + k + 1] - sS[g * (D + 1) + k];
            if (pr) val *= pr[k];
            yrow[rev ? D - 1 - k : k] = (T)val;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-(b, view) bin tables
// ---------------------------------------------------------------------------
template <typename TIn>
__global__ __launch_bounds__(256) void k_bin_tables(DevPlan P, const TIn* __restrict__ y, const TIn* __restrict__ m,
                                                    const double* __restrict__ pre, double scale,
                                                    double2* __restrict__ tabs, int nrows) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);  // b*V + v
    if (row >= nrows) return;                              // whole wave; no barriers below
    const int D = P.D;
    const int v = row % P.V;
    const bool rev = P.vp[v].rev;
    const double* e = P.edges + (size_t)v * (D + 1);
    const TIn* yr = y + (size_t)row * D;
    const TIn* mr = m ? m + (size_t)row * D : nullptr;
    const double* pr = pre ? pre + (size_t)v * D : nullptr;
    double2* tb = tabs + (size_t)row * (D + 2);
    // 64 consecutive bins per step: coalesced loads and stores, one wave scan per step
    double carry = 0.0;
    for (int k0 = 0; k0 < D; k0 += 64) {
        const int k = k0 + lane;
        double g = 0.0, ek = 0.0, wg = 0.0;
        if (k < D) {
            const int o = rev ? D - 1 - k : k;
            g = (double)yr[o];
            if (mr) g -= (double)mr[o];
            if (pr) g *= pr[k];
            g *= scale;
            ek = e[k];
            wg = (e[k + 1] - ek) * g;
        }
        const double inc = wave_scan_incl(wg, lane);
        if (k < D) tb[k + 1] = make_double2(carry + (inc - wg) - ek * g, g);
        carry += readlane_d(inc, 63);
    }
    if (lane == 0) {
        tb[0] = make_double2(0.0, 0.0);
        tb[D + 1] = make_double2(carry, 0.0);
    }
}

// ---------------------------------------------------------------------------
// Bin -> pixel (projection_t / backprojection)
// ---------------------------------------------------------------------------
enum { GEO_ED = 0, GEO_EA = 1, GEO_PARA = 2 };

template <int GEO>
__device__ __forceinline__ double eval_G(const double2* __restrict__ t, double a, double s, int p, const float (&lk)[4],
                                         float ea_sc, float ea_off, float hi) {
    const double pi = __builtin_fma(s, (double)p, a);
    const float pf = (float)pi;
    float q;
    if (GEO == GEO_PARA) {
        q = __builtin_fmaf(lk[0], pf, lk[1]);
    } else {
        const float num = __builtin_fmaf(lk[0], pf, lk[1]);
        const float den = __builtin_fmaf(lk[2], pf, lk[3]);
        q = num * __builtin_amdgcn_rcpf(den);
        if (GEO == GEO_EA) q = __builtin_fmaf(atanf(q), ea_sc, ea_off);
    }
    const double2 c = t[table_index(q, hi)];
    return __builtin_fma(pi, c.y, c.x);
}

// One launch covers both sweep axes: work items [0, B*ntx) are (b, tile of
// 4*LPW image rows) over the X-axis views, the rest (b, tile of columns) over
// the Y-axis views, written bottom-to-top to outYT.
//
// Per view: the next view's bin table is copied global -> LDS with
// global_load_lds_dwordx4 (asynchronous, no register staging) and the next
// view's LineCoefs of the tile are prefetched into one register per thread
// while the current view is evaluated; one barrier per view.
template <typename T, int LPW, int NPE, bool WEIGHTED, int GEO>
__global__ __launch_bounds__(256) void k_adj(DevPlan P, const double2* __restrict__ tabs, T* __restrict__ outX,
                                             T* __restrict__ outYT, int B, int ntx, int nty, double wscale) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int L = 4 * LPW;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nwx = B * ntx;
    int w = xcd_remap(blockIdx.x, gridDim.x, nwx + B * nty);
    if (w < 0) return;
    const int mode = w < nwx ? MODE_X : MODE_Y;
    if (mode == MODE_Y) w -= nwx;
    const int ntiles = mode == MODE_Y ? nty : ntx;
    const int b = w / ntiles, tile = w - b * ntiles;
    const int* vlist = mode == MODE_Y ? P.yviews : P.xviews;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const int nlines = mode == MODE_Y ? P.W : P.H;
    const int n = mode == MODE_Y ? P.H : P.W;
    const int D = P.D;
    const int npe = (n + 1 + 63) / 64;  // edge groups (<= NPE)
    const float hi = (float)(D + 1);
    const int tabn = D + 2;

    double2* stab = reinterpret_cast<double2*>(smem);                   // [2][tabn]
    double* scf = reinterpret_cast<double*>(stab + 2 * tabn);           // [2][L][LC_WORDS]

    const double2* tb_b = tabs + (size_t)b * P.V * tabn;
    // Asynchronous copy of view v's table into LDS buffer `buf`: wave w copies
    // entries [64*(w + 4r), 64*(w + 4r) + 64); lanes past the end are masked off.
    auto table_to_lds = [&](int v, int buf) {
        const double2* src = tb_b + (size_t)v * tabn;
        for (int j0 = wave * 64; j0 < tabn; j0 += NT) {
            if (j0 + lane < tabn)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(src + j0 + lane),
                    (__attribute__((address_space(3))) void*)(stab + buf * tabn + j0), 16, 0, 0);
        }
    };
    // LineCoef word `tid % 8` of tile line `tid / 8`, held by thread tid < L*8.
    const bool cthread = tid < L * LC_WORDS;
    const int cl = tile * L + tid / LC_WORDS;
    const bool cvalid = cthread && cl < nlines;
    auto coef_word = [&](int v) {
        return cvalid ? reinterpret_cast<const double*>(P.lines + (size_t)v * P.max_lines + cl)[tid % LC_WORDS]
                      : 0.0;
    };

    double acc[LPW][NPE];
#pragma unroll
    for (int q = 0; q < LPW; q++)
#pragma unroll
        for (int i = 0; i < NPE; i++) acc[q][i] = 0.0;

    table_to_lds(vlist[0], 0);
    if (cthread) scf[tid] = coef_word(vlist[0]);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    for (int iv = 0; iv < nv; iv++) {
        const int cur = iv & 1;
        const int v = vlist[iv];
        double cnext = 0.0;
        if (iv + 1 < nv) {
            const int vn = vlist[iv + 1];
            table_to_lds(vn, cur ^ 1);
            cnext = coef_word(vn);
        }
        const double2* t = stab + cur * tabn;
        const double* cf = scf + cur * L * LC_WORDS;
        const ViewParams& vp = P.vp[v];
        const float lk[4] = {vp.lk[0], vp.lk[1], vp.lk[2], vp.lk[3]};
        const float ea_sc = vp.ea_sc, ea_off = vp.ea_off;
#pragma unroll
        for (int q = 0; q < LPW; q++) {
            const int lq = q * 4 + wave;  // line within the tile
            const int l = tile * L + lq;
            if (l >= nlines) continue;    // wave-uniform
            const double a = cf[lq * LC_WORDS + LC_A0], s = cf[lq * LC_WORDS + LC_S];
            if (!WEIGHTED) {
#pragma unroll
                for (int i = 0; i < NPE; i++)
                    if (i < npe) acc[q][i] += eval_G<GEO>(t, a, s, lane + 64 * i, lk, ea_sc, ea_off, hi);
            } else {
                const double w0 = cf[lq * LC_WORDS + LC_W0], w1 = cf[lq * LC_WORDS + LC_W1];
                const double w2 = cf[lq * LC_WORDS + LC_W2];
                const double ws = cf[lq * LC_WORDS + LC_INVS] * wscale;
                auto wgt = [&](int p) {
                    const double u = __builtin_fma(w1, (double)p, w0);
                    const double d = __builtin_fma(u, u, w2);
                    return ws * (double)__builtin_amdgcn_rcpf((float)d);
                };
                double gprev = 0.0;
#pragma unroll
                for (int i = 0; i < NPE; i++) {
                    if (i < npe) {
                        const double gi = eval_G<GEO>(t, a, s, lane + 64 * i, lk, ea_sc, ea_off, hi);
                        if (i > 0) {
                            double right = shfl_down_d(gprev, 1);
                            const double r0 = readlane_d(gi, 0);
                            if (lane == 63) right = r0;
                            const int p = lane + 64 * (i - 1);
                            if (p < n) acc[q][i - 1] += wgt(p) * (right - gprev);
                        }
                        gprev = gi;
                    }
                }
                // pixels of the last edge group (their right edges are in the same group)
                {
                    const double right = shfl_down_d(gprev, 1);
                    const int p = lane + 64 * (npe - 1);
                    if (p < n) {
#pragma unroll
                        for (int i = 0; i < NPE; i++)
                            if (i == npe - 1) acc[q][i] += wgt(p) * (right - gprev);
                    }
                }
            }
        }
        if (iv + 1 < nv && cthread) scf[(cur ^ 1) * L * LC_WORDS + tid] = cnext;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }

    // Write pixel values.
#pragma unroll
    for (int q = 0; q < LPW; q++) {
        const int l = tile * L + q * 4 + wave;
        if (l >= nlines) continue;
        T* dst = mode == MODE_Y ? outYT + ((size_t)b * P.W + l) * P.H : outX + ((size_t)b * P.H + l) * P.W;
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            if (i >= npe) continue;  // wave-uniform
            const int p = lane + 64 * i;
            double val;
            if (!WEIGHTED) {
                double right = shfl_down_d(acc[q][i], 1);
                double r0 = 0.0;
#pragma unroll
                for (int k = 0; k < NPE; k++)
                    if (k == i + 1) r0 = readlane_d(acc[q][k], 0);
                if (lane == 63) right = r0;
                val = right - acc[q][i];
            } else {
                val = acc[q][i];
            }
            if (p < n) dst[p] = (T)val;
        }
    }
}

// ---------------------------------------------------------------------------
// out[b][r][c] (+)= yt[b][c][H-1-r]
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_combine(T* __restrict__ out, const T* __restrict__ yt, int H, int W, int hasX,
                                                 int hasY) {
    __shared__ T tile[32][33];
    const int b = blockIdx.z;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const T* yb = yt + (size_t)b * H * W;
    T* ob = out + (size_t)b * H * W;
    if (hasY) {
        // tile[cc][k] = yt[c0+cc][H-1-(r0+k)]
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

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

hipError_t launch_to_ylines(const void* x, void* xT, int dtype, int B, int H, int W, hipStream_t s) {
    if (B <= 0) return hipSuccess;
    dim3 grid((W + 31) / 32, (H + 31) / 32, B);
    if (dtype == DT_F32)
        hipLaunchKernelGGL(k_to_ylines<float>, grid, dim3(256), 0, s, (const float*)x, (float*)xT, H, W);
    else
        hipLaunchKernelGGL(k_to_ylines<double>, grid, dim3(256), 0, s, (const double*)x, (double*)xT, H, W);
    return hipGetLastError();
}

size_t fwd_lds_bytes(const DevPlan& P, bool weighted, int G) {
    const int n = P.nx ? P.W : P.H;
    const int nmax = P.ny && P.H > n ? P.H : n;
    const int ntab = weighted ? G : 1;
    const size_t tabs = (size_t)2 * ntab * (nmax + 2) * sizeof(double2) + (size_t)2 * G * LC_WORDS * sizeof(double) +
                        (size_t)2 * ntab * 4 * sizeof(double);
    const size_t ex = (size_t)G * (P.D + 1) * sizeof(double);
    return tabs > ex ? tabs : ex;
}

namespace {

template <typename T, int G, int NE, int M, bool WTD>
hipError_t fwd_launch_t(const DevPlan& P, const void* srcx, const void* srcy, void* y, int B, const double* post,
                        hipStream_t s) {
    const int ngx = (P.nx + G - 1) / G, ngy = (P.ny + G - 1) / G;
    const int nwork = (ngx + ngy) * B;
    const size_t lds = fwd_lds_bytes(P, WTD, G);
    auto kern = k_fwd<T, G, NE, M, WTD>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid_for(nwork)), dim3(NT), lds, s, P, (const T*)srcx, (const T*)srcy, (T*)y, B,
                       ngx, ngy, post);
    return hipGetLastError();
}

template <typename T, int G, int NE, bool WTD>
hipError_t fwd_m(const DevPlan& P, const void* sx, const void* sy, void* y, int B, const double* post, hipStream_t s) {
    int n = P.nx ? P.W : 0;
    if (P.ny && P.H > n) n = P.H;
    const int m = (n + NT - 1) / NT;
    if (m <= 1) return fwd_launch_t<T, G, NE, 1, WTD>(P, sx, sy, y, B, post, s);
    if (m <= 2) return fwd_launch_t<T, G, NE, 2, WTD>(P, sx, sy, y, B, post, s);
    if (m <= 4) return fwd_launch_t<T, G, NE, 4, WTD>(P, sx, sy, y, B, post, s);
    if (m <= 8) return fwd_launch_t<T, G, NE, 8, WTD>(P, sx, sy, y, B, post, s);
    return hipErrorInvalidValue;
}

template <typename T, int G, bool WTD>
hipError_t fwd_ne(const DevPlan& P, const void* sx, const void* sy, void* y, int B, const double* post,
                  hipStream_t s) {
    const int ne = (P.D + 1 + NT - 1) / NT;
    switch (ne) {
    case 1: return fwd_m<T, G, 1, WTD>(P, sx, sy, y, B, post, s);
    case 2: return fwd_m<T, G, 2, WTD>(P, sx, sy, y, B, post, s);
    case 3: return fwd_m<T, G, 3, WTD>(P, sx, sy, y, B, post, s);
    case 4: return fwd_m<T, G, 4, WTD>(P, sx, sy, y, B, post, s);
    case 5:
    case 6: return fwd_m<T, G, 6, WTD>(P, sx, sy, y, B, post, s);
    case 7:
    case 8: return fwd_m<T, G, 8, WTD>(P, sx, sy, y, B, post, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t fwd_dispatch(const DevPlan& P, const void* sx, const void* sy, void* y, int B, bool weighted,
                        const double* post, int G, hipStream_t s) {
    if (weighted) return fwd_ne<T, 2, true>(P, sx, sy, y, B, post, s);
    switch (G) {
    case 2: return fwd_ne<T, 2, false>(P, sx, sy, y, B, post, s);
    case 8: return fwd_ne<T, 8, false>(P, sx, sy, y, B, post, s);
    default: return fwd_ne<T, 4, false>(P, sx, sy, y, B, post, s);
    }
}

}  // namespace

hipError_t launch_fwd(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, int B, bool weighted,
                      const double* post, FwdShape shape, hipStream_t s) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if ((P.nx && P.W > NT * 8) || (P.ny && P.H > NT * 8)) return hipErrorInvalidValue;
    if (dtype == DT_F32) return fwd_dispatch<float>(P, srcx, srcy, y, B, weighted, post, shape.G, s);
    return fwd_dispatch<double>(P, srcx, srcy, y, B, weighted, post, shape.G, s);
}

hipError_t launch_bin_tables(const DevPlan& P, int in_dtype, const void* y, const void* m, const double* pre,
                             double scale, double2* tabs, int B, hipStream_t s) {
    const int nrows = B * P.V;
    if (nrows <= 0) return hipSuccess;
    dim3 grid((nrows + 3) / 4);
    if (in_dtype == DT_F32)
        hipLaunchKernelGGL(k_bin_tables<float>, grid, dim3(256), 0, s, P, (const float*)y, (const float*)m, pre, scale,
                           tabs, nrows);
    else
        hipLaunchKernelGGL(k_bin_tables<double>, grid, dim3(256), 0, s, P, (const double*)y, (const double*)m, pre,
                           scale, tabs, nrows);
    return hipGetLastError();
}

namespace {

template <typename T, int LPW, int NPE, bool WTD, int GEO>
hipError_t adj_launch_t(const DevPlan& P, const double2* tabs, void* outX, void* outYT, int B, double wscale,
                        hipStream_t s) {
    const int L = 4 * LPW;
    const int ntx = P.nx ? (P.H + L - 1) / L : 0;
    const int nty = P.ny ? (P.W + L - 1) / L : 0;
    const int nwork = (ntx + nty) * B;
    const size_t lds = (size_t)2 * (P.D + 2) * sizeof(double2) + (size_t)2 * L * LC_WORDS * sizeof(double);
    auto kern = k_adj<T, LPW, NPE, WTD, GEO>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid_for(nwork)), dim3(NT), lds, s, P, tabs, (T*)outX, (T*)outYT, B, ntx, nty,
                       wscale);
    return hipGetLastError();
}

template <typename T, bool WTD, int GEO>
hipError_t adj_npe(const DevPlan& P, const double2* tabs, void* outX, void* outYT, int B, double wscale,
                   hipStream_t s) {
    int n = P.nx ? P.W : 0;
    if (P.ny && P.H > n) n = P.H;
    const int npe = (n + 1 + 63) / 64;
    if (npe <= 5) return adj_launch_t<T, 8, 5, WTD, GEO>(P, tabs, outX, outYT, B, wscale, s);
    if (npe <= 9) return adj_launch_t<T, 4, 9, WTD, GEO>(P, tabs, outX, outYT, B, wscale, s);
    if (npe <= 17) return adj_launch_t<T, 2, 17, WTD, GEO>(P, tabs, outX, outYT, B, wscale, s);
    if (npe <= 33) return adj_launch_t<T, 1, 33, WTD, GEO>(P, tabs, outX, outYT, B, wscale, s);
    return hipErrorInvalidValue;
}

template <typename T, bool WTD>
hipError_t adj_geo(const DevPlan& P, const double2* tabs, void* outX, void* outYT, int B, double wscale,
                   hipStream_t s) {
    switch (P.scan_type) {
    case GEO_ED: return adj_npe<T, WTD, GEO_ED>(P, tabs, outX, outYT, B, wscale, s);
    case GEO_EA: return adj_npe<T, WTD, GEO_EA>(P, tabs, outX, outYT, B, wscale, s);
    default: return adj_npe<T, WTD, GEO_PARA>(P, tabs, outX, outYT, B, wscale, s);
    }
}

}  // namespace

hipError_t launch_adj(const DevPlan& P, int dtype, const double2* tabs, void* outX, void* outYT, int B, bool weighted,
                      double wscale, AdjShape, hipStream_t s) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if (dtype == DT_F32) {
        return weighted ? adj_geo<float, true>(P, tabs, outX, outYT, B, wscale, s)
                        : adj_geo<float, false>(P, tabs, outX, outYT, B, wscale, s);
    }
    return weighted ? adj_geo<double, true>(P, tabs, outX, outYT, B, wscale, s)
                    : adj_geo<double, false>(P, tabs, outX, outYT, B, wscale, s);
}

hipError_t launch_combine(void* out, const void* yt, int dtype, int B, int H, int W, bool hasX, bool hasY,
                          hipStream_t s) {
    if (B <= 0) return hipSuccess;
    dim3 grid((W + 31) / 32, (H + 31) / 32, B);
    if (dtype == DT_F32)
        hipLaunchKernelGGL(k_combine<float>, grid, dim3(256), 0, s, (float*)out, (const float*)yt, H, W, (int)hasX,
                           (int)hasY);
    else
        hipLaunchKernelGGL(k_combine<double>, grid, dim3(256), 0, s, (double*)out, (const double*)yt, H, W, (int)hasX,
                           (int)hasY);
    return hipGetLastError();
}

}  // namespace ctlib
