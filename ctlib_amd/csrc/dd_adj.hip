// Stand-alone launches of the bin -> pixel kernels (gfx950): projection_t and
// backprojection (k_adj) and the per-view backprojection_sv (k_bprj_sv).  The
// per-work-item device code lives in dd_device.h; kept apart from the line ->
// bin launches so the two translation units compile in parallel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd_device.h"
#include "kernels.h"

namespace ctlib {

using namespace dd;

namespace {

// One launch covers both sweep axes: work items [0, B*ntx) are (b, tile of
// 4*LPW image rows) over the X-axis views, the rest (b, tile of columns) over
// the Y-axis views, written bottom-to-top to outYT.  With nch > 1 every item is
// split into nch chunks of its axis's views; chunk k writes its partial image
// k*B*H*W elements into outX / outYT.
template <typename T, int LPW, int NPE, bool WEIGHTED, int GEO, bool EXACT>
__global__ __launch_bounds__(256) void k_adj(DevPlan P, const double2* __restrict__ tabs, T* __restrict__ outX,
                                             T* __restrict__ outYT, int B, int ntx, int nty, double wscale, int nch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nwx = B * ntx;
    int w = xcd_remap(blockIdx.x, gridDim.x, (nwx + B * nty) * nch);
    if (w < 0) return;
    const int ch = w % nch;
    w /= nch;
    const int mode = w < nwx ? MODE_X : MODE_Y;
    if (mode == MODE_Y) w -= nwx;
    const int ntiles = mode == MODE_Y ? nty : ntx;
    const int b = w / ntiles, tile = w - b * ntiles;
    const int nv = mode == MODE_Y ? P.ny : P.nx;
    const size_t part = (size_t)ch * B * P.H * P.W;
    adj_block<T, LPW, NPE, WEIGHTED, GEO, EXACT>(P, tabs, outX + part, outYT + part, b, mode, tile, wscale, smem,
                                                 ch * nv / nch, (ch + 1) * nv / nch);
}

template <typename T>
__global__ __launch_bounds__(256) void k_combine_parts(T* __restrict__ out, const T* __restrict__ xs,
                                                       const T* __restrict__ ys, int nch, int H, int W, int hasX,
                                                       int hasY) {
    __shared__ T tile[32][33];
    const int b = blockIdx.z;
    const int B = gridDim.z;
    const size_t img = (size_t)H * W;
    combine_parts_tile<T>(out + b * img, xs + b * img, ys + b * img, nch, (size_t)B * img, H, W, blockIdx.y * 32,
                          blockIdx.x * 32, hasX, hasY, tile);
}

// backprojection_sv: one view's term of the weighted backprojection, written to
// its own image, out[b][v][r][c] (fan_ed_kernel.cu:945-1161).  One block per
// (b, view); the view's bin table is staged in LDS.  X-axis views: a wave owns
// an image row and its lanes the pixel edges of 64 consecutive columns (NPE
// groups, all evaluated before the stores; the right edge comes from the
// neighbouring lane), so stores are row-contiguous.  Y-axis views: each lane
// owns a column and walks its pixels bottom to top, SV_UNROLL at a time, so
// the 64 lanes of a wave store 64 adjacent columns of one row.  The kernel is
// VALU-bound (95 % issue at B=8, 736 views, 512^2: its 6.2 GB of stores alone
// would take ~1.07 ms at the 5.7 TB/s a plain store kernel reaches), so the
// fp32 path evaluates the composed lookup (plan table adj_lk) and fp32 pixel
// weights, and steps each column's fp64 edge coordinate and float pixel
// counter by adds.
constexpr int SV_UNROLL = 8;

template <typename T, int GEO, int NPE>
__global__ __launch_bounds__(256) void k_bprj_sv(DevPlan P, const double2* __restrict__ tabs, T* __restrict__ out,
                                                 int nitems, double wscale) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int item = xcd_remap(blockIdx.x, gridDim.x, nitems);  // b * V + v
    if (item < 0) return;
    const int v = item % P.V;
    const int D = P.D, H = P.H, W = P.W;
    const int tabn = D + 2;
    double2* t = reinterpret_cast<double2*>(smem);
    const double2* src = tabs + (size_t)item * tabn;
    for (int j = threadIdx.x; j < tabn; j += NT) t[j] = src[j];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const ViewParams& vp = P.vp[v];
    const Lookup lu = view_lookup(P, v);
    const float hi = (float)(D + 1);
    const LineCoef* lc = P.lines + (size_t)v * P.max_lines;
    T* img = out + (size_t)item * H * W;
    auto wgt = [&](const LineCoef& c, int p) {
        const double u = __builtin_fma(c.w1, (double)p, c.w0);
        const double d = __builtin_fma(u, u, c.w2);
        return c.inv_s * wscale * (sizeof(T) == 8 ? rcp_refined(d) : (double)__builtin_amdgcn_rcpf((float)d));
    };
    if constexpr (sizeof(T) == 4) {
        // fp32 data: the lookup composed with each line's pixel-edge map (plan table adj_lk) and the
        // pixel weight in fp32, as in the bin->pixel kernels; the fp64 edge differences are rounded to
        // fp32 before the weight is applied.
        const float4* alk = P.adj_lk + (size_t)v * P.max_lines;
        // pixel weights from the plan's pre-scaled coefficients {64 c w1, c w1, c w0, c^2 w2}
        // (weight inv_s / ((w0 + w1 p)^2 + w2) = 1 / (u^2 + c^2 w2), u = c (w0 + w1 p)), with the
        // call's scale folded in per line: u / sqrt(wscale) and c^2 w2 / wscale
        const float4* awk = P.adj_wk + (size_t)v * P.max_lines;
        const float rs = (float)(1.0 / sqrt(wscale)), rs2 = (float)(1.0 / wscale);  // wscale > 0
        if (vp.xmode) {
            for (int r = wave; r < H; r += NT / 64) {
                const LineCoef c = lc[r];
                const LineLk lk = line_lk(c.a, c.s, alk[r], lane);
                T* dst = img + (size_t)r * W;
                double g[NPE];
                double pir = lk.pi0;
#pragma unroll
                for (int i = 0; i < NPE; i++) {
                    if (i > 0) pir += lk.s64;
                    g[i] = eval_G_lk<GEO>(t, pir, lk, i, lu, hi);
                }
                const float4 wk = awk[r];
                const float w1x64 = wk.x * rs, w1s = wk.y * rs, w2s = wk.w * rs2;
                const float u0 = __builtin_fmaf(w1s, (float)lane, wk.z * rs);
                // right edge of pixel lane + 64 i: edge lane + 1 of group i, for lane 63 edge 0 of
                // group i + 1 - both from the groups' rotations by one lane
                const bool last = lane == 63;
                double rcur = rot_next_d(g[0]);
#pragma unroll
                for (int i = 0; i < NPE; i++) {
                    // lane 63 of the last group is past the row: 64 * NPE > W
                    double right = rcur;
                    if (i + 1 < NPE) {
                        const double rnext = rot_next_d(g[i + 1]);
                        if (last) right = rnext;
                        rcur = rnext;
                    }
                    const int p = lane + 64 * i;
                    const float u = i == 0 ? u0 : __builtin_fmaf((float)i, w1x64, u0);
                    const float w = __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, w2s));
                    if (p < W) dst[p] = (float)(right - g[i]) * w;
                }
            }
        } else {
            for (int c0 = wave * 64; c0 < W; c0 += NT) {
                const int col = c0 + lane;
                const bool ok = col < W;
                const LineCoef c = lc[ok ? col : W - 1];
                const float4 ck = alk[ok ? col : W - 1];
                // edge p of the column: axis coordinate a + s p (stepped in fp64), lookup argument
                // (c0 p + c1) / (c2 p + c3) with p held exactly as a float counter
                const float cx = ck.x * (1.0f / 64.0f), cz = ck.z * (1.0f / 64.0f);
                const float4 wk = awk[ok ? col : W - 1];
                const float w0s = wk.z * rs, w1s = wk.y * rs, w2s = wk.w * rs2;
                auto G_edge = [&](double pi, float pf) {
                    const float num = __builtin_fmaf(cx, pf, ck.y);
                    float q;
                    if (GEO == GEO_PARA) {
                        q = num;
                    } else {
                        q = num * __builtin_amdgcn_rcpf(__builtin_fmaf(cz, pf, ck.w));
                        if (GEO == GEO_EA) q = __builtin_fmaf(atanf(q), lu.ea_sc, lu.ea_off);
                        if (GEO == GEO_EAP) {
                            const float tc = __builtin_amdgcn_fmed3f(q, -lu.T, lu.T);
                            const float s2 = tc * tc;
                            float z = lu.c[5];
#pragma unroll
                            for (int j = 4; j >= 0; j--) z = __builtin_fmaf(z, s2, lu.c[j]);
                            q = __builtin_fmaf(tc * z, lu.ea_sc, lu.ea_off);
                        }
                    }
                    const double2 e = t[table_index(q, hi)];
                    return __builtin_fma(pi, e.y, e.x);
                };
                double pi = c.a;
                float pf = 0.0f;
                double gprev = G_edge(pi, pf);
                T* dst = img + (size_t)(H - 1) * W + col;
#pragma unroll 8
                for (int i = 0; i < H; i++) {  // pixel i between edges i and i + 1
                    const float pix = pf;      // == i
                    pi += c.s;
                    pf += 1.0f;
                    const double gi = G_edge(pi, pf);
                    const float u = __builtin_fmaf(w1s, pix, w0s);
                    const float w = __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, w2s));
                    if (ok) dst[-(ptrdiff_t)i * W] = (float)(gi - gprev) * w;
                    gprev = gi;
                }
            }
        }
        return;
    }
    if (vp.xmode) {
        for (int r = wave; r < H; r += NT / 64) {
            const LineCoef c = lc[r];
            T* dst = img + (size_t)r * W;
            double g[NPE];
#pragma unroll
            for (int i = 0; i < NPE; i++) g[i] = eval_G<GEO, sizeof(T) == 8>(t, c.a, c.s, lane + 64 * i, lu, hi);
#pragma unroll
            for (int i = 0; i < NPE; i++) {
                double right = shfl_next_d(g[i]);
                if (i + 1 < NPE) {
                    const double r0 = readlane_d(g[i + 1], 0);
                    if (lane == 63) right = r0;
                }
                // lane 63 of the last group is past the row: 64 * NPE > W
                const int p = lane + 64 * i;
                if (p < W) dst[p] = (T)(wgt(c, p) * (right - g[i]));
            }
        }
    } else {
        for (int c0 = wave * 64; c0 < W; c0 += NT) {
            const int col = c0 + lane;
            const bool ok = col < W;
            const LineCoef c = lc[ok ? col : W - 1];
            double gprev = eval_G<GEO, sizeof(T) == 8>(t, c.a, c.s, 0, lu, hi);
            T* dst = img + (size_t)(H - 1) * W + col;
            int i = 0;
            for (; i + SV_UNROLL <= H; i += SV_UNROLL) {
                double g[SV_UNROLL];
#pragma unroll
                for (int u = 0; u < SV_UNROLL; u++) g[u] = eval_G<GEO, sizeof(T) == 8>(t, c.a, c.s, i + u + 1, lu, hi);
#pragma unroll
                for (int u = 0; u < SV_UNROLL; u++) {
                    const double left = u ? g[u - 1] : gprev;
                    if (ok) dst[-(ptrdiff_t)(i + u) * W] = (T)(wgt(c, i + u) * (g[u] - left));
                }
                gprev = g[SV_UNROLL - 1];
            }
            for (; i < H; i++) {
                const double gi = eval_G<GEO, sizeof(T) == 8>(t, c.a, c.s, i + 1, lu, hi);
                if (ok) dst[-(ptrdiff_t)i * W] = (T)(wgt(c, i) * (gi - gprev));
                gprev = gi;
            }
        }
    }
}
}  // namespace

namespace {

template <typename T, int LPW, int NPE, bool WTD, int GEO, bool EXACT>
hipError_t adj_launch_t(const DevPlan& P, const double2* tabs, void* outX, void* outYT, int B, double wscale,
                        int nch, hipStream_t s) {
    const int L = 4 * LPW;
    const int ntx = P.nx ? (P.H + L - 1) / L : 0;
    const int nty = P.ny ? (P.W + L - 1) / L : 0;
    const long long nw = (long long)(ntx + nty) * B * nch;
    if (nw > (1LL << 30)) return hipErrorInvalidValue;  // grid size (one block per work item)
    const int nwork = (int)nw;
    const size_t lds = adj_block_lds(P.D, L);
    auto kern = k_adj<T, LPW, NPE, WTD, GEO, EXACT>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid_for(nwork)), dim3(NT), lds, s, P, tabs, (T*)outX, (T*)outYT, B, ntx, nty,
                       wscale, nch);
    return hipGetLastError();
}

template <typename T, bool WTD, int GEO>
hipError_t adj_npe(const DevPlan& P, const double2* tabs, void* outX, void* outYT, int B, double wscale, int nch,
                   hipStream_t s) {
    int n = P.nx ? P.W : 0;
    if (P.ny && P.H > n) n = P.H;
    const int npe = (n + 1 + 63) / 64;
    // every line (rows of X views, columns of Y views) exactly 64 * (NPE - 1) pixels
    auto exact = [&](int NPE) { return (P.nx == 0 || P.W == 64 * (NPE - 1)) && (P.ny == 0 || P.H == 64 * (NPE - 1)); };
    if (npe <= 5) return exact(5) ? adj_launch_t<T, 8, 5, WTD, GEO, true>(P, tabs, outX, outYT, B, wscale, nch, s)
                                  : adj_launch_t<T, 8, 5, WTD, GEO, false>(P, tabs, outX, outYT, B, wscale, nch, s);
    if (npe <= 7) return exact(7) ? adj_launch_t<T, 4, 7, WTD, GEO, true>(P, tabs, outX, outYT, B, wscale, nch, s)
                                  : adj_launch_t<T, 4, 7, WTD, GEO, false>(P, tabs, outX, outYT, B, wscale, nch, s);
    if (npe <= 9) return exact(9) ? adj_launch_t<T, 4, 9, WTD, GEO, true>(P, tabs, outX, outYT, B, wscale, nch, s)
                                  : adj_launch_t<T, 4, 9, WTD, GEO, false>(P, tabs, outX, outYT, B, wscale, nch, s);
    if (npe <= 13) return exact(13) ? adj_launch_t<T, 2, 13, WTD, GEO, true>(P, tabs, outX, outYT, B, wscale, nch, s)
                                    : adj_launch_t<T, 2, 13, WTD, GEO, false>(P, tabs, outX, outYT, B, wscale, nch, s);
    if (npe <= 17) return exact(17) ? adj_launch_t<T, 2, 17, WTD, GEO, true>(P, tabs, outX, outYT, B, wscale, nch, s)
                                    : adj_launch_t<T, 2, 17, WTD, GEO, false>(P, tabs, outX, outYT, B, wscale, nch, s);
    if (npe <= 33) return adj_launch_t<T, 1, 33, WTD, GEO, false>(P, tabs, outX, outYT, B, wscale, nch, s);
    return hipErrorInvalidValue;
}

template <typename T, bool WTD>
hipError_t adj_geo(const DevPlan& P, const double2* tabs, void* outX, void* outYT, int B, double wscale, int nch,
                   hipStream_t s) {
    switch (P.scan_type) {
    case GEO_ED: return adj_npe<T, WTD, GEO_ED>(P, tabs, outX, outYT, B, wscale, nch, s);
    case GEO_EA:
        return P.ea_poly ? adj_npe<T, WTD, GEO_EAP>(P, tabs, outX, outYT, B, wscale, nch, s)
                         : adj_npe<T, WTD, GEO_EA>(P, tabs, outX, outYT, B, wscale, nch, s);
    default: return adj_npe<T, WTD, GEO_PARA>(P, tabs, outX, outYT, B, wscale, nch, s);
    }
}

}  // namespace

int adj_tile_lines(const DevPlan& P) {
    int n = P.nx ? P.W : 0;
    if (P.ny && P.H > n) n = P.H;
    const int npe = (n + 1 + 63) / 64;
    return npe <= 5 ? 32 : npe <= 9 ? 16 : npe <= 17 ? 8 : 4;
}

hipError_t launch_combine_parts(void* out, const void* xs, const void* ys, int dtype, int B, int H, int W, int nch,
                                bool hasX, bool hasY, hipStream_t s) {
    if (B <= 0) return hipSuccess;
    dim3 grid((W + 31) / 32, (H + 31) / 32, B);
    if (dtype == DT_F32)
        hipLaunchKernelGGL(k_combine_parts<float>, grid, dim3(256), 0, s, (float*)out, (const float*)xs,
                           (const float*)ys, nch, H, W, (int)hasX, (int)hasY);
    else
        hipLaunchKernelGGL(k_combine_parts<double>, grid, dim3(256), 0, s, (double*)out, (const double*)xs,
                           (const double*)ys, nch, H, W, (int)hasX, (int)hasY);
    return hipGetLastError();
}

hipError_t launch_adj(const DevPlan& P, int dtype, const double2* tabs, void* outX, void* outYT, int B, bool weighted,
                      double wscale, AdjShape sh, hipStream_t s) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if (sh.nch < 1) sh.nch = 1;
    if ((P.nx && sh.nch > P.nx) || (P.ny && sh.nch > P.ny)) return hipErrorInvalidValue;
    if (dtype == DT_F32) {
        return weighted ? adj_geo<float, true>(P, tabs, outX, outYT, B, wscale, sh.nch, s)
                        : adj_geo<float, false>(P, tabs, outX, outYT, B, wscale, sh.nch, s);
    }
    return weighted ? adj_geo<double, true>(P, tabs, outX, outYT, B, wscale, sh.nch, s)
                    : adj_geo<double, false>(P, tabs, outX, outYT, B, wscale, sh.nch, s);
}

namespace {

template <typename T, int GEO, int NPE>
hipError_t sv_launch_npe(const DevPlan& P, const double2* tabs, void* out, int B, double wscale, hipStream_t s) {
    const int nitems = B * P.V;
    const size_t lds = (size_t)(P.D + 2) * sizeof(double2);
    hipLaunchKernelGGL((k_bprj_sv<T, GEO, NPE>), dim3(grid_for(nitems)), dim3(NT), lds, s, P, tabs, (T*)out, nitems,
                       wscale);
    return hipGetLastError();
}

// Edge groups of an X-axis row: 64 * NPE >= W + 1.
template <typename T, int GEO>
hipError_t sv_launch_t(const DevPlan& P, const double2* tabs, void* out, int B, double wscale, hipStream_t s) {
    const int npe = (P.W + 1 + 63) / 64;
    if (npe <= 5) return sv_launch_npe<T, GEO, 5>(P, tabs, out, B, wscale, s);
    if (npe <= 9) return sv_launch_npe<T, GEO, 9>(P, tabs, out, B, wscale, s);
    if (npe <= 17) return sv_launch_npe<T, GEO, 17>(P, tabs, out, B, wscale, s);
    if (npe <= 33) return sv_launch_npe<T, GEO, 33>(P, tabs, out, B, wscale, s);
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t sv_geo(const DevPlan& P, const double2* tabs, void* out, int B, double wscale, hipStream_t s) {
    switch (P.scan_type) {
    case GEO_ED: return sv_launch_t<T, GEO_ED>(P, tabs, out, B, wscale, s);
    case GEO_EA:
        return P.ea_poly ? sv_launch_t<T, GEO_EAP>(P, tabs, out, B, wscale, s)
                         : sv_launch_t<T, GEO_EA>(P, tabs, out, B, wscale, s);
    default: return sv_launch_t<T, GEO_PARA>(P, tabs, out, B, wscale, s);
    }
}

}  // namespace

hipError_t launch_bprj_sv(const DevPlan& P, int dtype, const double2* tabs, void* out, int B, double wscale,
                          hipStream_t s) {
    if (B <= 0) return hipSuccess;
    if (!(wscale > 0.0)) return hipErrorInvalidValue;  // the fp32 path folds 1/sqrt(wscale) into its weights
    if ((size_t)B * P.V > (size_t)INT32_MAX || (size_t)(P.D + 2) * sizeof(double2) > 64 * 1024)
        return hipErrorInvalidValue;
    if (dtype == DT_F32) return sv_geo<float>(P, tabs, out, B, wscale, s);
    return sv_geo<double>(P, tabs, out, B, wscale, s);
}

}  // namespace ctlib
