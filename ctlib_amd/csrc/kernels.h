// Device-side plan view and kernel launchers shared by the C ABI (abi.cpp) and
// the HIP kernels (dd_kernels.hip, ramp_mfma.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "plan.h"

struct ctlib_plan;

namespace ctlib {

// Device pointers into one plan allocation (all read-only after upload).
struct DevPlan {
    const ViewParams* vp;     // [V]
    const double* edges;      // [V][D+1]
    const double* post_prj;   // [V][D]
    const double* post_bt;    // [V][D]
    const LineCoef* lines;    // [V][max_lines]
    const int* xviews;        // [nx]
    const int* yviews;        // [ny]
    const double* taps;       // [2D-1]
    const double* preweight;  // [D]
    const double* lookup_d;   // [V][6] lookup coefficients in fp64 (used for fp64 data)
    const float4* adj_lk;     // [V][max_lines] lookup composed with the line's pixel-edge map (HostPlan)
    const float4* adj_wk;     // [V][max_lines] pre-scaled pixel-weight coefficients (HostPlan)
    int nx, ny;
    int V, D, W, H, max_lines;
    int scan_type;
    double dAng;
    float ea_c[6];            // fan-EA polynomial atan (see HostPlan)
    float ea_T;
    int ea_poly;
};

enum { DT_F32 = 0, DT_F64 = 1 };
enum { MODE_X = 0, MODE_Y = 1 };

// Launch-shape choices the autotuner may change (see tune.cpp).
struct FwdShape {                 // line->bin kernel
    int G;                        // views per block
    int nch;                      // line chunks (<= 1: none)
};
struct AdjShape { int nch; };     // view chunks of the bin->pixel kernel (<= 1: none)

// x [B][H][W] -> xT [B][W][H] with xT[b][c][i] = x[b][H-1-i][c]  (Y-mode lines, bottom to top)
hipError_t launch_to_ylines(const void* x, void* xT, int dtype, int B, int H, int W, hipStream_t s);

// Line -> bin operator (projection; weighted = backprojection_t), both sweep axes in one launch.
// `srcx` holds image rows ([B][H][W]), `srcy` the bottom-to-top column copy ([B][W][H]).
// With shape.nch = k > 1 each axis's image lines are split into k chunks (k <= lines of each
// axis) and chunk c writes its partial sinogram at y + c*B*V*D (elements); launch_sum_parts
// adds them up.  ltab (unweighted only; ignored when weighted): the line tables of
// launch_line_tables for the same srcx / srcy, copied into LDS line by line instead of built
// per block; null: each block builds the tables of its lines.
hipError_t launch_fwd(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, int B,
                      bool weighted, const double* post, FwdShape shape, hipStream_t s,
                      const double2* ltab = nullptr);
// Cumulative table of every image line of both axes (one wave per line) into ltab, of
// fwd_line_table_bytes(P, B): per slice the X-axis lines (rows of srcx) then the Y-axis lines
// (rows of srcy), 256*M + 2 entries {C_j, S_j} per line, M = fwd_px(P).
hipError_t launch_line_tables(const DevPlan& P, int dtype, const void* srcx, const void* srcy, double2* ltab, int B,
                              hipStream_t s);
// out[i] = sum_k parts[k*count + i], k < nch (out may equal parts)
hipError_t launch_sum_parts(void* out, const void* parts, int dtype, int nch, size_t count, hipStream_t s);
size_t fwd_lds_bytes(const DevPlan& P, bool weighted, int G);
// Pixels per thread M of the line -> bin launch (lines of at most 256*M pixels; 0: too long).
int fwd_px(const DevPlan& P);
// Bytes of the prebuilt line tables of a batch of B: B * (lines of both axes) * (256*M + 2) * 16.
size_t fwd_line_table_bytes(const DevPlan& P, int B);

// projection fused with the residual bin tables of projection_t: writes y = A x (y may be null)
// and tabs[b][v] = tables of g = (y - m) * post_prj, so that launch_adj yields A^T (A x - m);
// rowsq (may be null) receives sum_k (y - m)^2 per row [B][V], y rounded to the data type.
// ltab as for launch_fwd.
hipError_t launch_fwd_resid(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, const void* m,
                            double2* tabs, int B, FwdShape shape, hipStream_t s, double* rowsq = nullptr,
                            const double2* ltab = nullptr);

// Weighted line -> bin operator (backprojection_t; post may be null) with wave-private line
// tables, for lines of at most 1024 pixels (fwd_w_supported); launch_fwd dispatches to it.
bool fwd_w_supported(const DevPlan& P);
hipError_t launch_fwd_w(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, int B,
                        const double* post, int nch, hipStream_t s);

// Per (b, view) cumulative-integral tables {C_j, g_j}, j = 0..D+1, of g_k = (y - m)[o(k)] * pre[k] * scale;
// with m and rowsq not null also rowsq[b][v] = sum_k (y - m)_k^2 (y rounded to the data type).
// nparts > 1: y holds nparts partial sinograms pstride elements apart, summed in FP64 in order,
// and the sum is written to yout (rounded to the data type) when yout is not null.
hipError_t launch_bin_tables(const DevPlan& P, int in_dtype, const void* y, const void* m, const double* pre,
                             double scale, double2* tabs, int B, hipStream_t s, double* rowsq = nullptr,
                             int nparts = 1, size_t pstride = 0, void* yout = nullptr);

// Bin -> pixel operator (projection_t; weighted = backprojection, pixel weight * wscale), both axes in
// one launch.  X lines are written to outX [B][H][W]; Y lines to outYT [B][W][H] (bottom-to-top per column).
// With shape.nch = k > 1 each axis's views are split into k chunks (k <= views of each axis) and chunk c
// writes its partial images at outX + c*B*H*W and outYT + c*B*H*W (elements); launch_combine_parts sums them.
hipError_t launch_adj(const DevPlan& P, int dtype, const double2* tabs, void* outX, void* outYT, int B,
                      bool weighted, double wscale, AdjShape shape, hipStream_t s);
// Image lines per work item of launch_adj (work items = B * (ceil(H/L) [X views] + ceil(W/L) [Y views])).
int adj_tile_lines(const DevPlan& P);
// out[b][r][c] = sum_k xs[k][b][r][c] (hasX) + sum_k ys[k][b][c][H-1-r] (hasY), k < nch
hipError_t launch_combine_parts(void* out, const void* xs, const void* ys, int dtype, int B, int H, int W, int nch,
                                bool hasX, bool hasY, hipStream_t s);

// Per-view weighted backprojection (backprojection_sv): out[b][v][H][W] = view v's term of
// backprojection, pixel weight * wscale (wscale > 0), from the bin tables of launch_bin_tables.
hipError_t launch_bprj_sv(const DevPlan& P, int dtype, const double2* tabs, void* out, int B, double wscale,
                          hipStream_t s);

// out[b][r][c] = (hasX ? out[b][r][c] : 0) + (hasY ? yt[b][c][H-1-r] : 0)
hipError_t launch_combine(void* out, const void* yt, int dtype, int B, int H, int W, bool hasX, bool hasY,
                          hipStream_t s);

// Ramp filtering of `rows` sinogram rows as a dense Toeplitz GEMM on FP64 MFMA:
//   out[r][j] = scale * post[j] * sum_m (pre[m] * in[r][m]) * h(m - j)     (pre/post may be null = 1)
// with h the RL taps of the plan.  in/out dtypes are DT_F32 or DT_F64.
hipError_t launch_ramp(const DevPlan& P, int in_dtype, const void* in, int out_dtype, void* out, long rows,
                       const double* pre, const double* post, double scale, hipStream_t s);

// kNN graph weights (laplacian): cols [n][k] int64, vals [n][k] float32 from points x [n][d] (dtype DT_F32 or
// DT_F64 as CTLIB_F32 / CTLIB_F64); ws of laplacian_ws_bytes.  Supported (d, k): laplacian_supported.
bool laplacian_supported(int d, int k, int dtype);
size_t laplacian_ws_bytes(int64_t n, int d, int k, int dtype);
hipError_t launch_laplacian(const void* x, int64_t n, int d, int k, int dtype, int64_t* cols, float* vals, void* ws,
                            hipStream_t s);

// Device view of a plan (for the batch module).
DevPlan plan_dev(const struct ::ctlib_plan* plan);

}  // namespace ctlib
