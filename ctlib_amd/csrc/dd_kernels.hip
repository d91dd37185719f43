// Stand-alone launches of the line -> bin kernels (gfx950): projection,
// backprojection_t and the residual-fused forward (k_fwd), the per-(slice,
// view) bin tables of the adjoints, and the Y-axis line copy / combine
// transposes.  The per-work-item device code lives in dd_device.h; the
// persistent heterogeneous batch kernel (batch.hip) runs the same block
// functions, and the bin -> pixel launches live in dd_adj.hip.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd_device.h"
#include "kernels.h"

namespace ctlib {

using namespace dd;

namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_to_ylines(const T* __restrict__ x, T* __restrict__ xT, int H, int W) {
    __shared__ T tile[32][33];
    const int b = blockIdx.z;
    ylines_tile<T>(x + (size_t)b * H * W, xT + (size_t)b * H * W, H, W, blockIdx.y * 32, blockIdx.x * 32, tile);
}

// fp32 variants with 24 edge slots per lane and up to 4 pixels per thread
// (e.g. 1536 bins over 1024-pixel lines) are held to 168 VGPRs: three waves per
// SIMD instead of two, without spills.
template <typename T, int NEM, int M, bool WEIGHTED>
constexpr int fwd_waves() { return sizeof(T) == 4 && !WEIGHTED && NEM > 20 && NEM <= 24 && M <= 4 ? 3 : 1; }

// Line tables of every line of a batch into `lt`: per slice the X-axis lines (rows of
// srcx, n = W) then the Y-axis lines (rows of srcy, n = H), one wave per line.
template <typename T, int M>
__global__ __launch_bounds__(256) void k_line_tables(DevPlan P, const T* __restrict__ srcx, const T* __restrict__ srcy,
                                                     double2* __restrict__ lt, int B) {
    const int hx = P.nx ? P.H : 0, wy = P.ny ? P.W : 0;
    const int nl = hx + wy;
    const int line = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (line >= B * nl) return;  // whole wave
    const int b = line / nl, l = line - b * nl;
    const T* row = l < hx ? srcx + ((size_t)b * P.H + l) * P.W : srcy + ((size_t)b * P.W + (l - hx)) * P.H;
    line_table_wave<T, M>(row, l < hx ? P.W : P.H, lt + (size_t)line * ltab_n<M>(), threadIdx.x & 63);
}

// Line-table buffers of k_fwd_lt (one line ahead).
constexpr int FWD_LT_BUFFERS = 2;

// k_fwd over the prebuilt line tables of k_line_tables (unweighted).
template <typename T, int VPW, int NEM, int M, bool RESID, bool EDGEX>
__global__ __launch_bounds__(256, (fwd_waves<T, NEM, M, false>())) void k_fwd_lt(
    DevPlan P, const double2* __restrict__ lt, T* __restrict__ y, int B, int ngx, int ngy,
    const double* __restrict__ post, const T* __restrict__ m, double2* __restrict__ tabs, int nch,
    double* __restrict__ rowsq) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nwx = B * ngx;
    int w = xcd_remap(blockIdx.x, gridDim.x, (nwx + B * ngy) * nch);
    if (w < 0) return;
    const int ch = w % nch;
    w /= nch;
    const int mode = w < nwx ? MODE_X : MODE_Y;
    if (mode == MODE_Y) w -= nwx;
    const int ngroups = mode == MODE_Y ? ngy : ngx;
    const int b = w / ngroups, grp = w - b * ngroups;
    const int nl = mode == MODE_Y ? P.W : P.H;
    const int hx = P.nx ? P.H : 0, wy = P.ny ? P.W : 0;
    const double2* lts = lt + ((size_t)b * (hx + wy) + (mode == MODE_Y ? hx : 0)) * ltab_n<M>();
    fwd_block_lt<T, VPW, NEM, M, RESID, EDGEX, FWD_LT_BUFFERS>(P, lts, y + (size_t)ch * B * P.V * P.D, b, mode, grp, post,
                                                          m, tabs, smem, ch * nl / nch, (ch + 1) * nl / nch, rowsq);
}

// One launch covers both sweep axes: work items [0, B*ngx) are (b, group of
// 4*VPW X-axis views) and read image rows from `srcx`; the rest are Y-axis
// groups reading the bottom-to-top column copy `srcy`.
template <typename T, int VPW, int NEM, int M, bool WEIGHTED, bool RESID, bool FULL, bool EDGEX>
__global__ __launch_bounds__(256, (fwd_waves<T, NEM, M, WEIGHTED>())) void k_fwd(DevPlan P, const T* __restrict__ srcx, const T* __restrict__ srcy,
                                             T* __restrict__ y, int B, int ngx, int ngy,
                                             const double* __restrict__ post, const T* __restrict__ m,
                                             double2* __restrict__ tabs, int nch, double* __restrict__ rowsq) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nwx = B * ngx;
    int w = xcd_remap(blockIdx.x, gridDim.x, (nwx + B * ngy) * nch);
    if (w < 0) return;
    const int ch = w % nch;
    w /= nch;
    const int mode = w < nwx ? MODE_X : MODE_Y;
    if (mode == MODE_Y) w -= nwx;
    const int ngroups = mode == MODE_Y ? ngy : ngx;
    const int b = w / ngroups, grp = w - b * ngroups;
    const int nl = mode == MODE_Y ? P.W : P.H;
    fwd_block<T, VPW, NEM, M, WEIGHTED, RESID, FULL, EDGEX>(P, mode == MODE_Y ? srcy : srcx,
                                                            y + (size_t)ch * B * P.V * P.D, b, mode, grp, post, m,
                                                            tabs, smem, ch * nl / nch, (ch + 1) * nl / nch, rowsq);
}

// out[i] = sum_k parts[k * count + i], k < nch (partial sinograms of a line-chunked launch);
// out may be the first partial (each element is read and written by the same thread).
template <typename T>
__global__ __launch_bounds__(256) void k_sum_parts(T* out, const T* parts, int nch, size_t count) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) {
        T v = parts[i];
        for (int k = 1; k < nch; k++) v += parts[(size_t)k * count + i];
        out[i] = v;
    }
}

// Per-(b, view) bin tables {C_j, g_j} of g_k = ((y - m)[o(k)] * pre[k]) * scale, j = 0..D+1;
// with m and rowsq not null also rowsq[row] = sum_k (y - m)_k^2.  nparts > 1: y is the sum of
// nparts partial sinograms pstride apart (FP64, in order), also stored to yout if not null;
// rowsq then uses the sum rounded to TIn, as a stored y would be.
template <typename TIn>
__global__ __launch_bounds__(256) void k_bin_tables(DevPlan P, const TIn* __restrict__ y, const TIn* __restrict__ m,
                                                    const double* __restrict__ pre, double scale,
                                                    double2* __restrict__ tabs, int nrows,
                                                    double* __restrict__ rowsq, int nparts, size_t pstride,
                                                    TIn* __restrict__ yout) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);  // b*V + v
    if (row >= nrows) return;                              // whole wave; no barriers below
    const int D = P.D;
    const int v = row % P.V;
    const bool rev = P.vp[v].rev;
    const double* e = P.edges + (size_t)v * (D + 1);
    const TIn* yr = y + (size_t)row * D;
    TIn* yo = yout ? yout + (size_t)row * D : nullptr;
    const TIn* mr = m ? m + (size_t)row * D : nullptr;
    const double* pr = pre ? pre + (size_t)v * D : nullptr;
    double2* tb = tabs + (size_t)row * (D + 2);
    // 64 consecutive bins per step: coalesced loads and stores, one wave scan per step
    // (issuing four steps' loads before their scans measured slower)
    double carry = 0.0, rsq = 0.0;
    for (int k0 = 0; k0 < D; k0 += 64) {
        const int k = k0 + lane;
        double g = 0.0, ek = 0.0, wg = 0.0;
        if (k < D) {
            const int o = rev ? D - 1 - k : k;
            g = (double)yr[o];
            double r = 0.0;
            if (nparts > 1) {
                for (int c = 1; c < nparts; c++) g += (double)yr[c * pstride + o];
                if (yo) yo[o] = (TIn)g;
                if (mr) r = (double)(TIn)g - (double)mr[o];
            }
            if (mr) {
                g -= (double)mr[o];
                if (nparts <= 1) r = g;
            }
            rsq = __builtin_fma(r, r, rsq);
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
    if (rowsq && mr) {
        const double tot = readlane_d(wave_scan_incl(rsq, lane), 63);
        if (lane == 0) rowsq[row] = tot;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_combine(T* __restrict__ out, const T* __restrict__ yt, int H, int W, int hasX,
                                                 int hasY) {
    __shared__ T tile[32][33];
    const int b = blockIdx.z;
    combine_tile<T>(out + (size_t)b * H * W, yt + (size_t)b * H * W, H, W, blockIdx.y * 32, blockIdx.x * 32, hasX, hasY,
                    tile);
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

int fwd_px(const DevPlan& P) {
    int n = P.nx ? P.W : 0;
    if (P.ny && P.H > n) n = P.H;
    const int m = (n + NT - 1) / NT;
    if (m <= 1 && (P.nx == 0 || P.W == NT) && (P.ny == 0 || P.H == NT)) return 1;
    return m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : 0;
}

size_t fwd_line_table_bytes(const DevPlan& P, int B) {
    const int M = fwd_px(P);
    if (B <= 0 || M == 0 || P.nx + P.ny == 0) return 0;
    const size_t nlt = (size_t)(P.nx ? P.H : 0) + (P.ny ? P.W : 0);
    return (size_t)B * nlt * (NT * M + 2) * sizeof(double2);
}

size_t fwd_lds_bytes(const DevPlan& P, bool weighted, int G) {
    const int n = P.nx ? P.W : P.H;
    const int nmax = P.ny && P.H > n ? P.H : n;
    return fwd_block_lds(nmax, G, weighted, 16);
}

namespace {

struct FwdArgs {
    const void* srcx;
    const void* srcy;
    void* y;
    int B;
    const double* post;
    const void* m;
    double2* tabs;
    int nch;
    double* rowsq;
    const double2* lt;  // line tables of launch_line_tables (unweighted), or null: each block builds its own
};

template <typename T, int VPW, int NEM, int M, bool WTD, bool RESID, bool FULL, bool EDGEX>
hipError_t fwd_launch_t(const DevPlan& P, const FwdArgs& a, hipStream_t s) {
    constexpr int G = 4 * VPW;
    const int ngx = (P.nx + G - 1) / G, ngy = (P.ny + G - 1) / G;
    const long long nw = (long long)(ngx + ngy) * a.B * a.nch;
    if (nw > (1LL << 30)) return hipErrorInvalidValue;  // grid size (one block per work item)
    const int nwork = (int)nw;
    if constexpr (!WTD) {
        if (a.lt) {
            const size_t lds = fwd_lt_lds<M>(G, FWD_LT_BUFFERS);
            auto kern = k_fwd_lt<T, VPW, NEM, M, RESID, EDGEX>;
            if (lds > 64 * 1024) {
                hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(kern, dim3(grid_for(nwork)), dim3(NT), lds, s, P, (const double2*)a.lt, (T*)a.y, a.B,
                               ngx, ngy, a.post, (const T*)a.m, a.tabs, a.nch, a.rowsq);
            return hipGetLastError();
        }
    }
    const size_t lds = fwd_lds_bytes(P, WTD, G);
    auto kern = k_fwd<T, VPW, NEM, M, WTD, RESID, FULL, EDGEX>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid_for(nwork)), dim3(NT), lds, s, P, (const T*)a.srcx, (const T*)a.srcy, (T*)a.y,
                       a.B, ngx, ngy, a.post, (const T*)a.m, a.tabs, a.nch, a.rowsq);
    return hipGetLastError();
}

// Every swept line (rows of X views, columns of Y views) exactly NT * M pixels.
bool fwd_full(const DevPlan& P, int M) { return (P.nx == 0 || P.W == NT * M) && (P.ny == 0 || P.H == NT * M); }

template <typename T, int VPW, int NEM, bool WTD, bool RESID, bool EDGEX = false>
hipError_t fwd_m(const DevPlan& P, const FwdArgs& a, hipStream_t s) {
    switch (fwd_px(P)) {
    case 1: return fwd_launch_t<T, VPW, NEM, 1, WTD, RESID, true, EDGEX>(P, a, s);
    case 2: return fwd_full(P, 2) ? fwd_launch_t<T, VPW, NEM, 2, WTD, RESID, true, EDGEX>(P, a, s)
                                  : fwd_launch_t<T, VPW, NEM, 2, WTD, RESID, false, EDGEX>(P, a, s);
    case 4: return fwd_full(P, 4) ? fwd_launch_t<T, VPW, NEM, 4, WTD, RESID, true, EDGEX>(P, a, s)
                                  : fwd_launch_t<T, VPW, NEM, 4, WTD, RESID, false, EDGEX>(P, a, s);
    case 8: return fwd_launch_t<T, VPW, NEM, 8, WTD, RESID, false, EDGEX>(P, a, s);
    default: return hipErrorInvalidValue;
    }
}

// Edge slots of 64 per wave and view: NEM >= ceil((D+1)/64); two views per wave only up to NEM 12.
// D = 512, 1024 or 1536 bins (unweighted): the last edge D is evaluated packed (EDGEX).
template <typename T, int VPW, bool WTD, bool RESID>
hipError_t fwd_ne(const DevPlan& P, const FwdArgs& a, hipStream_t s) {
    const int ne = (P.D + 1 + 63) / 64;
    if (!WTD && P.D == 512) return fwd_m<T, VPW, 8, WTD, RESID, true>(P, a, s);
    if (!WTD && VPW == 1 && P.D == 1024) return fwd_m<T, 1, 16, WTD, RESID, true>(P, a, s);
    if (!WTD && VPW == 1 && P.D == 1536) return fwd_m<T, 1, 24, WTD, RESID, true>(P, a, s);
    if (ne <= 4) return fwd_m<T, VPW, 4, WTD, RESID>(P, a, s);
    if (ne <= 6) return fwd_m<T, VPW, 6, WTD, RESID>(P, a, s);
    if (ne <= 8) return fwd_m<T, VPW, 8, WTD, RESID>(P, a, s);
    if (ne <= 12) return fwd_m<T, VPW, 12, WTD, RESID>(P, a, s);
    if (VPW != 1) return hipErrorInvalidValue;
    if (ne <= 14) return fwd_m<T, 1, 14, WTD, RESID>(P, a, s);
    if (ne <= 16) return fwd_m<T, 1, 16, WTD, RESID>(P, a, s);
    if (ne <= 20) return fwd_m<T, 1, 20, WTD, RESID>(P, a, s);
    if (ne <= 24) return fwd_m<T, 1, 24, WTD, RESID>(P, a, s);
    if (ne <= 28) return fwd_m<T, 1, 28, WTD, RESID>(P, a, s);
    if (ne <= 32) return fwd_m<T, 1, 32, WTD, RESID>(P, a, s);
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t fwd_dispatch(const DevPlan& P, const FwdArgs& a, bool weighted, bool resid, int G, hipStream_t s) {
    const bool two = G == 8 && (P.D + 1 + 63) / 64 <= 12;
    if (weighted) return fwd_ne<T, 1, true, false>(P, a, s);
    if (resid) return two ? fwd_ne<T, 2, false, true>(P, a, s) : fwd_ne<T, 1, false, true>(P, a, s);
    return two ? fwd_ne<T, 2, false, false>(P, a, s) : fwd_ne<T, 1, false, false>(P, a, s);
}

template <typename T>
hipError_t line_tables_t(const DevPlan& P, const void* srcx, const void* srcy, double2* lt, int B, hipStream_t s) {
    const long long ntab = (long long)B * ((P.nx ? P.H : 0) + (P.ny ? P.W : 0));
    if (ntab > (1LL << 31) - 4) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((ntab + 3) / 4));
    switch (fwd_px(P)) {
    case 1: hipLaunchKernelGGL((k_line_tables<T, 1>), grid, dim3(256), 0, s, P, (const T*)srcx, (const T*)srcy, lt, B); break;
    case 2: hipLaunchKernelGGL((k_line_tables<T, 2>), grid, dim3(256), 0, s, P, (const T*)srcx, (const T*)srcy, lt, B); break;
    case 4: hipLaunchKernelGGL((k_line_tables<T, 4>), grid, dim3(256), 0, s, P, (const T*)srcx, (const T*)srcy, lt, B); break;
    case 8: hipLaunchKernelGGL((k_line_tables<T, 8>), grid, dim3(256), 0, s, P, (const T*)srcx, (const T*)srcy, lt, B); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_line_tables(const DevPlan& P, int dtype, const void* srcx, const void* srcy, double2* ltab, int B,
                              hipStream_t s) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if (dtype == DT_F32) return line_tables_t<float>(P, srcx, srcy, ltab, B, s);
    return line_tables_t<double>(P, srcx, srcy, ltab, B, s);
}

hipError_t launch_fwd(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, int B, bool weighted,
                      const double* post, FwdShape shape, hipStream_t s, const double2* ltab) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if ((P.nx && P.W > NT * 8) || (P.ny && P.H > NT * 8) || P.D > 2046) return hipErrorInvalidValue;
    const int nch = shape.nch < 1 ? 1 : shape.nch;
    if ((P.nx && nch > P.H) || (P.ny && nch > P.W)) return hipErrorInvalidValue;
    if (weighted && fwd_w_supported(P)) return launch_fwd_w(P, dtype, srcx, srcy, y, B, post, nch, s);
    FwdArgs a{srcx, srcy, y, B, post, nullptr, nullptr, nch, nullptr, weighted ? nullptr : ltab};
    if (dtype == DT_F32) return fwd_dispatch<float>(P, a, weighted, false, shape.G, s);
    return fwd_dispatch<double>(P, a, weighted, false, shape.G, s);
}

hipError_t launch_fwd_resid(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, const void* m,
                            double2* tabs, int B, FwdShape shape, hipStream_t s, double* rowsq, const double2* ltab) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if ((P.nx && P.W > NT * 8) || (P.ny && P.H > NT * 8) || P.D > 2046) return hipErrorInvalidValue;
    FwdArgs a{srcx, srcy, y, B, P.post_prj, m, tabs, 1, rowsq, ltab};
    if (dtype == DT_F32) return fwd_dispatch<float>(P, a, false, true, shape.G, s);
    return fwd_dispatch<double>(P, a, false, true, shape.G, s);
}

hipError_t launch_sum_parts(void* out, const void* parts, int dtype, int nch, size_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (dtype == DT_F32)
        hipLaunchKernelGGL(k_sum_parts<float>, dim3((unsigned)blocks), dim3(256), 0, s, (float*)out,
                           (const float*)parts, nch, count);
    else
        hipLaunchKernelGGL(k_sum_parts<double>, dim3((unsigned)blocks), dim3(256), 0, s, (double*)out,
                           (const double*)parts, nch, count);
    return hipGetLastError();
}

hipError_t launch_bin_tables(const DevPlan& P, int in_dtype, const void* y, const void* m, const double* pre,
                             double scale, double2* tabs, int B, hipStream_t s, double* rowsq, int nparts,
                             size_t pstride, void* yout) {
    const int nrows = B * P.V;
    if (nrows <= 0) return hipSuccess;
    dim3 grid((nrows + 3) / 4);
    if (in_dtype == DT_F32)
        hipLaunchKernelGGL(k_bin_tables<float>, grid, dim3(256), 0, s, P, (const float*)y, (const float*)m, pre, scale,
                           tabs, nrows, rowsq, nparts, pstride, (float*)yout);
    else
        hipLaunchKernelGGL(k_bin_tables<double>, grid, dim3(256), 0, s, P, (const double*)y, (const double*)m, pre,
                           scale, tabs, nrows, rowsq, nparts, pstride, (double*)yout);
    return hipGetLastError();
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
