// C ABI of libctlib_amd.so (declared in include/ctlib_abi.h).
//
// Replaces the reference's dispatch layer (ctlib.cpp:15-99) and its per-geometry
// host wrappers (fan_ed_kernel.cu:1163-1382, fan_ea_kernel.cu:923-1107,
// para_kernel.cu:685-849).  Differences by design:
//   * options are parsed once into an immutable device-resident plan;
//   * ctlib_run launches on the caller's stream, allocates nothing and never
//     synchronises (the reference makes 6-12 .item() device syncs per call);
//   * outputs are written exactly once (the reference zero-fills and then
//     accumulates with global atomics);
//   * invalid input returns a status code (the reference exit(0)s on an unknown
//     scan_type, ctlib.cpp:26).
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ctlib_abi.h"
#include "kernels.h"
#include "plan.h"

using namespace ctlib;

struct ctlib_plan {
    HostPlan host;
    int device = 0;
    void* dmem = nullptr;
    DevPlan dev{};
    // Launch-shape overrides; each call reads them once (Tuning) so that the workspace it
    // checks and the launches it makes agree even if another thread changes them.
    std::atomic<int> fwd_G{0};    // autotuned views per line->bin block (0 = automatic)
    std::atomic<int> adj_nch{0};  // view chunks of the bin->pixel kernel (0 = automatic)
    std::atomic<int> fwd_nch{0};  // line chunks of the line->bin kernel (0 = automatic)
    std::atomic<int> fwd_lt{0};   // line->bin line tables: 0 prebuilt in the workspace, 1 per block
};

namespace ctlib {
DevPlan plan_dev(const ::ctlib_plan* p) { return p->dev; }
}  // namespace ctlib

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& why) {
    g_err = why;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return CTLIB_ERR_HIP;
}

// roctx range around one C-ABI call, so that a marker trace (rocprofv3
// --marker-trace) groups each call's kernels under the operator's name.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
};

const char* op_range_name(int op) {
    static const char* names[] = {"ctlib:projection", "ctlib:projection_t", "ctlib:backprojection",
                                  "ctlib:backprojection_t", "ctlib:backprojection_t_true", "ctlib:fbp",
                                  "ctlib:fbp_t", "ctlib:backprojection_sv"};
    return op >= 0 && op < (int)(sizeof(names) / sizeof(names[0])) ? names[op] : "ctlib:op";
}

// Restores the calling thread's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) return;
        if (prev != dev && hipSetDevice(dev) != hipSuccess) return;
        ok = true;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// The plan's launch-shape overrides as one call sees them.
struct Tuning {
    int fwd_G, adj_nch, fwd_nch, fwd_lt;
};

Tuning snap(const ctlib_plan* p) {
    return Tuning{p->fwd_G.load(std::memory_order_relaxed), p->adj_nch.load(std::memory_order_relaxed),
                  p->fwd_nch.load(std::memory_order_relaxed), p->fwd_lt.load(std::memory_order_relaxed)};
}

// Views per block of the line -> bin launch.  Weighted: 4.  Unweighted, with two views per
// wave possible (<= 12 edge slots of 64 per view): `tuned` (4 or 8) if set, else 4 over
// prebuilt line tables and 8 over tables built per block (measured at fan-ED 512^2,
// 736 views, B=32: 2.27 ms at 4 vs 2.40 at 8 prebuilt; 2.84 at 8 vs 3.40 at 4 per block).
int fwd_views(const DevPlan& P, bool weighted, int tuned, bool prebuilt) {
    if (weighted || (P.D + 1 + 63) / 64 > 12) return 4;
    if (tuned == 4 || tuned == 8) return tuned;
    return prebuilt ? 4 : 8;
}

// ---- scratch of ctlib_run_noalloc: one fixed device array per GPU, no allocation per call.
// `scratch` names the host-side stand-in of the array; its device address on a GPU comes
// from hipGetSymbolAddress with that GPU current (resolve_scratch, called when a plan is
// created on the device) and is the only form in which host code passes it on.
constexpr size_t kScratchBytes = CTLIB_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
std::mutex g_scr_mu;
void* g_scr_ptr[kMaxDevices] = {};
bool g_scr_tried[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[CTLIB_NOALLOC_SCRATCH_BYTES];

namespace {

// With `device` current: records this device's address of `scratch` (once per device).
void resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scr_mu);
    if (g_scr_tried[device]) return;
    g_scr_tried[device] = true;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scr_ptr[device] = ptr;
    else (void)hipGetLastError();
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scr_mu);
    return g_scr_ptr[device];
}

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

size_t dsize(int dtype) { return dtype == CTLIB_F64 ? sizeof(double) : sizeof(float); }

void to_geom(const ctlib_geom* in, Geom* g) {
    g->views = in->views;
    g->dets = in->dets;
    g->width = in->width;
    g->height = in->height;
    g->scan_type = in->scan_type;
    g->dImg = in->dImg;
    g->dDet = in->dDet;
    g->Ang0 = in->Ang0;
    g->dAng = in->dAng;
    g->s2r = in->s2r;
    g->d2r = in->d2r;
    g->binshift = in->binshift;
}

// View chunks of the bin -> pixel launch for a batch of B.  Each (image, line tile) work
// item's views are split and the partial images summed afterwards, so that the launch
// has about 4096 blocks: fewer leave CUs idle or badly balanced at the end (measured on
// 512^2 images: batch 1 runs 5x faster with 12 chunks, batch 32 5% faster with 2).
// A chunk costs a partial image of traffic and a fixed per-block set-up of about 16
// views' work, so chunks keep at least 16 views of each axis and there are at most 12.
int adj_chunks(const ctlib_plan* p, const Tuning& T, int64_t B) {
    const DevPlan& P = p->dev;
    int vmin = 1 << 30;
    if (P.nx) vmin = P.nx;
    if (P.ny && P.ny < vmin) vmin = P.ny;
    if (vmin == 1 << 30) return 1;
    if (T.adj_nch > 0) return T.adj_nch < vmin ? T.adj_nch : vmin;
    const int L = adj_tile_lines(P);
    const int64_t items = B * ((P.nx ? (P.H + L - 1) / L : 0) + (P.ny ? (P.W + L - 1) / L : 0));
    int64_t nch = (4096 + items - 1) / items;
    if (nch > 12) nch = 12;
    if (nch > vmin / 16) nch = vmin / 16;
    return nch < 1 ? 1 : (int)nch;
}

// Line chunks of the line -> bin launch for a batch of B, on the same grounds as
// adj_chunks: a (image, group of G views) work item's image lines are split so the
// launch has about 4096 blocks, each chunk writing a partial sinogram that is summed
// afterwards; chunks keep at least 32 lines of each axis and there are at most 16.
int fwd_chunks(const ctlib_plan* p, const Tuning& T, int64_t B, bool weighted) {
    const DevPlan& P = p->dev;
    int lmin = 1 << 30;
    if (P.nx) lmin = P.H;
    if (P.ny && P.W < lmin) lmin = P.W;
    if (lmin == 1 << 30) return 1;
    if (T.fwd_nch > 0) return T.fwd_nch < lmin ? T.fwd_nch : lmin;
    const int G = fwd_views(P, weighted, T.fwd_G, !weighted && T.fwd_lt == 0);
    const int64_t items = B * ((P.nx + G - 1) / G + (P.ny + G - 1) / G);
    int64_t nch = (4096 + items - 1) / items;
    if (nch > 16) nch = 16;
    if (nch > lmin / 32) nch = lmin / 32;
    return nch < 1 ? 1 : (int)nch;
}

// Scratch layout of every op (byte offsets into ws).
struct WsLayout {
    size_t ylines = 0, tables = 0, yt = 0, xp = 0, sp = 0, rows = 0, ltab = 0, total = 0;
    size_t n_ylines = 0, n_tables = 0, n_yt = 0, n_xp = 0, n_sp = 0, n_rows = 0, n_ltab = 0;
    int nch = 1;   // view chunks of the bin -> pixel launch (partials in xp / yt when > 1)
    int fnch = 1;  // line chunks of the line -> bin launch (partial sinograms in sp when > 1)
    int fwd_G = 0; // views per line -> bin block override (Tuning::fwd_G)

    // prebuilt line tables of the unweighted line -> bin launch (null: built per block)
    double2* line_tables(char* ws) const { return n_ltab ? reinterpret_cast<double2*>(ws + ltab) : nullptr; }
};

WsLayout layout(const ctlib_plan* p, const Tuning& T, int op, int64_t B, int dtype) {
    const DevPlan& P = p->dev;
    const size_t es = dsize(dtype);
    const size_t img = (size_t)B * P.H * P.W;
    const size_t sino = (size_t)B * P.V * P.D;
    WsLayout L;
    L.fwd_G = T.fwd_G;
    auto sino_partials = [&](bool weighted) {
        L.fnch = fwd_chunks(p, T, B, weighted);
        L.n_sp = L.fnch > 1 ? L.fnch * sino * es : 0;
    };
    auto partials = [&]() {
        L.nch = adj_chunks(p, T, B);
        L.n_yt = P.ny ? L.nch * img * es : 0;
        L.n_xp = (L.nch > 1 && P.nx) ? L.nch * img * es : 0;
    };
    switch (op) {
    case CTLIB_OP_PRJ:
    case CTLIB_OP_BPRJ_T:
    case CTLIB_OP_BPRJ_T_TRUE:
        L.n_ylines = P.ny ? img * es : 0;
        sino_partials(op != CTLIB_OP_PRJ);
        if (op == CTLIB_OP_PRJ && T.fwd_lt == 0) L.n_ltab = fwd_line_table_bytes(P, (int)B);
        break;
    case CTLIB_OP_PRJ_T:
    case CTLIB_OP_BPRJ:
        L.n_tables = (size_t)B * P.V * (P.D + 2) * sizeof(double2);
        partials();
        break;
    case CTLIB_OP_FBP:
        L.n_rows = sino * sizeof(double);
        L.n_tables = (size_t)B * P.V * (P.D + 2) * sizeof(double2);
        partials();
        break;
    case CTLIB_OP_FBP_T:
        L.n_ylines = P.ny ? img * es : 0;
        L.n_rows = sino * es;
        sino_partials(true);
        break;
    case CTLIB_OP_BPRJ_SV:
        L.n_tables = (size_t)B * P.V * (P.D + 2) * sizeof(double2);
        break;
    default:
        break;
    }
    size_t off = 0;
    L.ylines = off; off += align_up(L.n_ylines);
    L.tables = off; off += align_up(L.n_tables);
    L.yt = off;     off += align_up(L.n_yt);
    L.xp = off;     off += align_up(L.n_xp);
    L.sp = off;     off += align_up(L.n_sp);
    L.rows = off;   off += align_up(L.n_rows);
    L.ltab = off;   off += align_up(L.n_ltab);
    L.total = off;
    return L;
}

// Launch shape of the line -> bin kernel under layout L.
FwdShape fwd_shape(const DevPlan& P, bool weighted, const WsLayout& L) {
    return FwdShape{fwd_views(P, weighted, L.fwd_G, !weighted && L.n_ltab > 0), L.fnch};
}

// ---- kernel timing (HIP event pairs around the main launches, on their stream)
enum { K_FWD = 0, K_ADJ = 1, K_RAMP = 2, K_AUX = 3, K_SV = 4 };

struct TimedLaunch {
    hipEvent_t a, b;
    int cls;
    double bytes, flops;
};

std::mutex g_tmu;
bool g_timing = false;
std::vector<TimedLaunch> g_rec;
std::vector<hipEvent_t> g_pool;

hipEvent_t pool_get() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Set while the calling thread runs ctlib_run_noalloc: its launches record no events
// (event creation may allocate).
thread_local bool t_no_timing = false;
struct NoTiming {
    NoTiming() { t_no_timing = true; }
    ~NoTiming() { t_no_timing = false; }
};

class KernelTimer {
  public:
    KernelTimer(int cls, double bytes, double flops, hipStream_t s) : s_(s) {
        if (t_no_timing) return;
        std::lock_guard<std::mutex> lock(g_tmu);
        if (!g_timing) return;
        t_ = TimedLaunch{pool_get(), pool_get(), cls, bytes, flops};
        on_ = t_.a && t_.b && hipEventRecord(t_.a, s_) == hipSuccess;
    }
    void done() {
        if (!on_) return;
        on_ = false;
        if (hipEventRecord(t_.b, s_) != hipSuccess) return;
        std::lock_guard<std::mutex> lock(g_tmu);
        g_rec.push_back(t_);
    }
    ~KernelTimer() { done(); }

  private:
    hipStream_t s_;
    TimedLaunch t_{};
    bool on_ = false;
};

// Algorithmic bytes of one DD operator application (sweep model): every view streams
// the whole image once and its own sinogram row, es * B * V * (H*W + D).
double sweep_bytes(const DevPlan& P, int B, int dt) {
    const double es = dt == DT_F64 ? 8.0 : 4.0;
    return es * B * (double)P.V * ((double)P.H * P.W + P.D);
}

int check_run_args(const ctlib_plan* p, int op, int dtype, const void* in, void* out, int64_t B) {
    if (!p) return fail(CTLIB_ERR_INVALID_ARG, "plan is NULL");
    if (op < CTLIB_OP_PRJ || op > CTLIB_OP_BPRJ_SV) return fail(CTLIB_ERR_INVALID_ARG, "unknown op code");
    if (dtype != CTLIB_F32 && dtype != CTLIB_F64) return fail(CTLIB_ERR_UNSUPPORTED_DTYPE, "dtype must be f32 or f64");
    if (B < 0 || B > (1 << 24)) return fail(CTLIB_ERR_INVALID_ARG, "batch out of range");
    if (B > 0 && (!in || !out)) return fail(CTLIB_ERR_INVALID_ARG, "in/out must be device pointers");
    return CTLIB_OK;
}

// `need` is the size of the layout the call will launch with (computed once per call).
int check_ws(size_t need, const void* ws, size_t ws_bytes, const char* what) {
    if (need > 0 && (!ws || ws_bytes < need))
        return fail(CTLIB_ERR_WORKSPACE, std::string("workspace smaller than ") + what);
    return CTLIB_OK;
}

hipError_t run_line_tables(const DevPlan& P, int dtype, const void* srcx, const void* srcy, double2* lt, int B,
                           hipStream_t s) {
    KernelTimer t(K_AUX, 0, 0, s);
    return launch_line_tables(P, dtype, srcx, srcy, lt, B, s);
}

// Sinogram-output ops: projection and backprojection_t.
hipError_t run_to_sino(const ctlib_plan* p, int dtype, const void* in, void* out, int B, bool weighted,
                       const double* post, const WsLayout& L, char* ws, hipStream_t s) {
    const DevPlan& P = p->dev;
    hipError_t e = hipSuccess;
    void* yl = nullptr;
    if (P.ny) {
        yl = ws + L.ylines;
        KernelTimer t(K_AUX, 0, 0, s);
        e = launch_to_ylines(in, yl, dtype, B, P.H, P.W, s);
        if (e != hipSuccess) return e;
    }
    void* dst = L.fnch > 1 ? ws + L.sp : out;
    double2* lt = weighted ? nullptr : L.line_tables(ws);
    if (lt) e = run_line_tables(P, dtype, in, yl, lt, B, s);
    if (e != hipSuccess) return e;
    {
        KernelTimer t(K_FWD, sweep_bytes(P, B, dtype), 0, s);
        e = launch_fwd(P, dtype, in, yl, dst, B, weighted, post, fwd_shape(P, weighted, L), s, lt);
    }
    if (e != hipSuccess || L.fnch <= 1) return e;
    KernelTimer t(K_AUX, 0, 0, s);
    return launch_sum_parts(out, dst, dtype, L.fnch, (size_t)B * P.V * P.D, s);
}

// Bin tables for the image-output ops.
hipError_t run_tables(const ctlib_plan* p, int in_dtype, const void* y, const void* m, const double* pre,
                      double2* tabs, int B, hipStream_t s, double* rowsq = nullptr) {
    KernelTimer t(K_AUX, 0, 0, s);
    return launch_bin_tables(p->dev, in_dtype, y, m, pre, 1.0, tabs, B, s, rowsq);
}

// Image-output ops from prepared bin tables: projection_t and backprojection.
hipError_t run_to_image(const ctlib_plan* p, int dtype, const double2* tabs, void* out, int B, bool weighted,
                        double wscale, const WsLayout& L, char* ws, hipStream_t s) {
    const DevPlan& P = p->dev;
    void* yt = ws + L.yt;
    void* outX = L.nch > 1 ? ws + L.xp : out;
    hipError_t e;
    {
        KernelTimer t(K_ADJ, sweep_bytes(P, B, dtype), 0, s);
        e = launch_adj(P, dtype, tabs, outX, yt, B, weighted, wscale, AdjShape{L.nch}, s);
    }
    if (e != hipSuccess) return e;
    KernelTimer t(K_AUX, 0, 0, s);
    if (L.nch > 1) return launch_combine_parts(out, outX, yt, dtype, B, P.H, P.W, L.nch, P.nx > 0, P.ny > 0, s);
    if (!P.ny) return e;
    return launch_combine(out, yt, dtype, B, P.H, P.W, P.nx > 0, true, s);
}

hipError_t run_ramp(const ctlib_plan* p, int in_dt, const void* in, int out_dt, void* out, int B, const double* pre,
                    const double* post, double scale, hipStream_t s) {
    const DevPlan& P = p->dev;
    const long rows = (long)B * P.V;
    KernelTimer t(K_RAMP, 0, 2.0 * rows * (double)P.D * P.D, s);
    return launch_ramp(P, in_dt, in, out_dt, out, rows, pre, post, scale, s);
}

// One operator with the scratch layout L at w (ctlib_run and ctlib_run_noalloc).
int run_with(const ctlib_plan* p, int op, int dtype, const void* in, void* out, int64_t batch, const WsLayout& L,
             char* w, hipStream_t s) {
    DeviceGuard guard(p->device);
    if (!guard.ok) return fail(CTLIB_ERR_HIP, "cannot select the plan's device");
    TraceRange range(op_range_name(op));
    const DevPlan& P = p->dev;
    const int B = (int)batch;
    const int dt = dtype == CTLIB_F64 ? DT_F64 : DT_F32;
    double2* tabs = reinterpret_cast<double2*>(w + L.tables);
    hipError_t e = hipSuccess;
    switch (op) {
    case CTLIB_OP_PRJ:
        e = run_to_sino(p, dt, in, out, B, false, P.post_prj, L, w, s);
        break;
    case CTLIB_OP_BPRJ_T:
        e = run_to_sino(p, dt, in, out, B, true, P.scan_type == SCAN_FAN_EA ? P.post_bt : nullptr, L, w, s);
        break;
    case CTLIB_OP_BPRJ_T_TRUE:
        e = run_to_sino(p, dt, in, out, B, true, nullptr, L, w, s);
        break;
    case CTLIB_OP_PRJ_T:
        e = run_tables(p, dt, in, nullptr, P.post_prj, tabs, B, s);
        if (e == hipSuccess) e = run_to_image(p, dt, tabs, out, B, false, 1.0, L, w, s);
        break;
    case CTLIB_OP_BPRJ:
        e = run_tables(p, dt, in, nullptr, nullptr, tabs, B, s);
        if (e == hipSuccess) e = run_to_image(p, dt, tabs, out, B, true, 1.0, L, w, s);
        break;
    case CTLIB_OP_FBP: {
        // pre-weight + RL ramp (Toeplitz GEMM) -> weighted backprojection * dAng/2 (fan_ed_kernel.cu:1349-1380)
        double* rows = reinterpret_cast<double*>(w + L.rows);
        e = run_ramp(p, dt, in, DT_F64, rows, B, P.preweight, nullptr, 1.0, s);
        if (e == hipSuccess) e = run_tables(p, DT_F64, rows, nullptr, nullptr, tabs, B, s);
        if (e == hipSuccess) e = run_to_image(p, dt, tabs, out, B, true, P.dAng / 2, L, w, s);
        break;
    }
    case CTLIB_OP_FBP_T: {
        // adjoint of fbp: exact backprojection adjoint, then ramp * dAng/2, then pre-weight
        void* rows = w + L.rows;
        e = run_to_sino(p, dt, in, rows, B, true, nullptr, L, w, s);
        if (e == hipSuccess) e = run_ramp(p, dt, rows, dt, out, B, nullptr, P.preweight, P.dAng / 2, s);
        break;
    }
    case CTLIB_OP_BPRJ_SV: {
        // per-view terms of backprojection, one image per (b, view) (fan_ed_kernel.cu:1299-1331)
        if ((size_t)B * P.V > (size_t)INT32_MAX) return fail(CTLIB_ERR_INVALID_ARG, "batch * views exceeds 2^31");
        e = run_tables(p, dt, in, nullptr, nullptr, tabs, B, s);
        if (e == hipSuccess) {
            KernelTimer t(K_SV, sweep_bytes(P, B, dt), 0, s);
            e = launch_bprj_sv(P, dt, tabs, out, B, 1.0, s);
        }
        break;
    }
    default:
        return fail(CTLIB_ERR_INVALID_ARG, "unknown op code");
    }
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return CTLIB_OK;
}

}  // namespace

extern "C" {

const char* ctlib_last_error(void) { return g_err.c_str(); }

const char* ctlib_status_string(int status) {
    switch (status) {
    case CTLIB_OK: return "ok";
    case CTLIB_ERR_INVALID_ARG: return "invalid argument";
    case CTLIB_ERR_BAD_SCAN_TYPE: return "bad scan_type";
    case CTLIB_ERR_SHAPE: return "shape mismatch";
    case CTLIB_ERR_HIP: return "HIP runtime error";
    case CTLIB_ERR_UNSUPPORTED_DTYPE: return "unsupported dtype";
    case CTLIB_ERR_WORKSPACE: return "workspace too small";
    case CTLIB_ERR_GEOMETRY: return "unsupported geometry";
    default: return "unknown status";
    }
}

int ctlib_parse_options(const double* opts, int n, ctlib_geom* out) {
    if (!out) return fail(CTLIB_ERR_INVALID_ARG, "output geometry is NULL");
    Geom g{};
    std::string why;
    int rc = parse_options(opts, n, &g, &why);
    if (rc) return fail(rc, why);
    out->views = g.views;
    out->dets = g.dets;
    out->width = g.width;
    out->height = g.height;
    out->scan_type = g.scan_type;
    out->dImg = g.dImg;
    out->dDet = g.dDet;
    out->Ang0 = g.Ang0;
    out->dAng = g.dAng;
    out->s2r = g.s2r;
    out->d2r = g.d2r;
    out->binshift = g.binshift;
    return CTLIB_OK;
}

int ctlib_plan_create(const ctlib_geom* geom, int device, ctlib_plan** out) {
    if (!geom || !out) return fail(CTLIB_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    Geom g{};
    to_geom(geom, &g);
    if (g.views < 1 || g.dets < 1 || g.width < 1 || g.height < 1)
        return fail(CTLIB_ERR_INVALID_ARG, "sizes must be positive");
    if (g.dets > 2046 || g.width > 2048 || g.height > 2048)
        return fail(CTLIB_ERR_GEOMETRY, "sizes above dets 2046 / width, height 2048 are not supported");
    ctlib_plan* p = new (std::nothrow) ctlib_plan();
    if (!p) return fail(CTLIB_ERR_INVALID_ARG, "out of host memory");
    std::string why;
    int rc = build_plan(g, &p->host, &why);
    if (rc) {
        delete p;
        return fail(rc, why);
    }
    p->device = device;
    const HostPlan& H = p->host;
    const int V = g.views, D = g.dets;
    // One device allocation holds every table.
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += align_up(bytes);
        return o;
    };
    const size_t o_vp = take(sizeof(ViewParams) * V);
    const size_t o_edges = take(sizeof(double) * V * (D + 1));
    const size_t o_pp = take(sizeof(double) * V * D);
    const size_t o_pb = take(sizeof(double) * V * D);
    const size_t o_lines = take(sizeof(LineCoef) * H.lines.size());
    const size_t o_xv = take(sizeof(int) * (H.xviews.size() + 1));
    const size_t o_yv = take(sizeof(int) * (H.yviews.size() + 1));
    const size_t o_taps = take(sizeof(double) * H.ramp_taps.size());
    const size_t o_pw = take(sizeof(double) * H.ramp_preweight.size());
    const size_t o_lkd = take(sizeof(double) * H.lookup_d.size());
    const size_t o_alk = take(sizeof(float) * H.adj_lk.size());
    const size_t o_awk = take(sizeof(float) * H.adj_wk.size());
    const size_t total = off;

    DeviceGuard guard(device);
    if (!guard.ok) {
        delete p;
        return fail(CTLIB_ERR_HIP, "cannot select device " + std::to_string(device));
    }
    hipError_t e = hipMalloc(&p->dmem, total);
    if (e != hipSuccess) {
        delete p;
        return hip_fail(e, "hipMalloc(plan)");
    }
    char* base = static_cast<char*>(p->dmem);
    struct Copy { size_t off; const void* src; size_t bytes; };
    const Copy copies[] = {
        {o_vp, H.vp.data(), sizeof(ViewParams) * V},
        {o_edges, H.edges.data(), sizeof(double) * V * (D + 1)},
        {o_pp, H.post_prj.data(), sizeof(double) * V * D},
        {o_pb, H.post_bprj_t.data(), sizeof(double) * V * D},
        {o_lines, H.lines.data(), sizeof(LineCoef) * H.lines.size()},
        {o_xv, H.xviews.data(), sizeof(int) * H.xviews.size()},
        {o_yv, H.yviews.data(), sizeof(int) * H.yviews.size()},
        {o_taps, H.ramp_taps.data(), sizeof(double) * H.ramp_taps.size()},
        {o_pw, H.ramp_preweight.data(), sizeof(double) * H.ramp_preweight.size()},
        {o_lkd, H.lookup_d.data(), sizeof(double) * H.lookup_d.size()},
        {o_alk, H.adj_lk.data(), sizeof(float) * H.adj_lk.size()},
        {o_awk, H.adj_wk.data(), sizeof(float) * H.adj_wk.size()},
    };
    for (const Copy& c : copies) {
        if (!c.bytes) continue;
        e = hipMemcpy(base + c.off, c.src, c.bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p->dmem);
            delete p;
            return hip_fail(e, "hipMemcpy(plan)");
        }
    }
    resolve_scratch(device);
    DevPlan& P = p->dev;
    P.vp = reinterpret_cast<const ViewParams*>(base + o_vp);
    P.edges = reinterpret_cast<const double*>(base + o_edges);
    P.post_prj = reinterpret_cast<const double*>(base + o_pp);
    P.post_bt = reinterpret_cast<const double*>(base + o_pb);
    P.lines = reinterpret_cast<const LineCoef*>(base + o_lines);
    P.xviews = reinterpret_cast<const int*>(base + o_xv);
    P.yviews = reinterpret_cast<const int*>(base + o_yv);
    P.taps = reinterpret_cast<const double*>(base + o_taps);
    P.preweight = reinterpret_cast<const double*>(base + o_pw);
    P.lookup_d = reinterpret_cast<const double*>(base + o_lkd);
    P.adj_lk = reinterpret_cast<const float4*>(base + o_alk);
    P.adj_wk = reinterpret_cast<const float4*>(base + o_awk);
    P.nx = (int)H.xviews.size();
    P.ny = (int)H.yviews.size();
    P.V = V;
    P.D = D;
    P.W = g.width;
    P.H = g.height;
    P.max_lines = H.max_lines;
    P.scan_type = g.scan_type;
    P.dAng = g.dAng;
    for (int k = 0; k < 6; k++) P.ea_c[k] = H.ea_c[k];
    P.ea_T = H.ea_T;
    P.ea_poly = H.ea_poly;
    *out = p;
    return CTLIB_OK;
}

void ctlib_plan_destroy(ctlib_plan* p) {
    if (!p) return;
    if (p->dmem) {
        DeviceGuard guard(p->device);
        (void)hipFree(p->dmem);
    }
    delete p;
}

size_t ctlib_workspace_bytes(const ctlib_plan* p, int op, int64_t batch, int dtype) {
    if (!p || batch <= 0) return 0;
    return layout(p, snap(p), op, batch, dtype).total;
}

int ctlib_plan_set_tuning(ctlib_plan* p, int fwd_views_per_block) {
    if (!p) return fail(CTLIB_ERR_INVALID_ARG, "NULL plan");
    if (fwd_views_per_block != 0 && fwd_views_per_block != 4 && fwd_views_per_block != 8)
        return fail(CTLIB_ERR_INVALID_ARG, "views per block must be 0 (automatic), 4 or 8");
    p->fwd_G = fwd_views_per_block;
    return CTLIB_OK;
}

int ctlib_plan_set_line_tables(ctlib_plan* p, int per_block) {
    if (!p) return fail(CTLIB_ERR_INVALID_ARG, "NULL plan");
    if (per_block != 0 && per_block != 1) return fail(CTLIB_ERR_INVALID_ARG, "per_block must be 0 or 1");
    p->fwd_lt = per_block;
    return CTLIB_OK;
}

int ctlib_plan_set_chunks(ctlib_plan* p, int line_chunks, int view_chunks) {
    if (!p) return fail(CTLIB_ERR_INVALID_ARG, "NULL plan");
    if (line_chunks < 0 || line_chunks > 64 || view_chunks < 0 || view_chunks > 64)
        return fail(CTLIB_ERR_INVALID_ARG, "chunk counts must be 0 (automatic) or 1..64");
    p->fwd_nch = line_chunks;
    p->adj_nch = view_chunks;
    return CTLIB_OK;
}

int ctlib_plan_geom(const ctlib_plan* p, ctlib_geom* out) {
    if (!p || !out) return fail(CTLIB_ERR_INVALID_ARG, "NULL argument");
    const Geom& g = p->host.g;
    out->views = g.views;
    out->dets = g.dets;
    out->width = g.width;
    out->height = g.height;
    out->scan_type = g.scan_type;
    out->dImg = g.dImg;
    out->dDet = g.dDet;
    out->Ang0 = g.Ang0;
    out->dAng = g.dAng;
    out->s2r = g.s2r;
    out->d2r = g.d2r;
    out->binshift = g.binshift;
    return CTLIB_OK;
}

int ctlib_plan_view_info(const ctlib_plan* p, int view, int* xmode, int* rev) {
    if (!p || !xmode || !rev) return fail(CTLIB_ERR_INVALID_ARG, "NULL argument");
    if (view < 0 || view >= p->host.g.views) return fail(CTLIB_ERR_INVALID_ARG, "view out of range");
    *xmode = p->host.vp[view].xmode;
    *rev = p->host.vp[view].rev;
    return CTLIB_OK;
}

int ctlib_plan_export(const ctlib_geom* geom, int table, void* dst, size_t* bytes) {
    if (!geom || !bytes) return fail(CTLIB_ERR_INVALID_ARG, "NULL argument");
    Geom g{};
    to_geom(geom, &g);
    HostPlan H;
    std::string why;
    int rc = build_plan(g, &H, &why);
    if (rc) return fail(rc, why);
    const void* src = nullptr;
    size_t n = 0;
    float ea[8];
    switch (table) {
    case 0: src = H.vp.data(); n = H.vp.size() * sizeof(ViewParams); break;
    case 1: src = H.edges.data(); n = H.edges.size() * sizeof(double); break;
    case 2: src = H.post_prj.data(); n = H.post_prj.size() * sizeof(double); break;
    case 3: src = H.post_bprj_t.data(); n = H.post_bprj_t.size() * sizeof(double); break;
    case 4: src = H.lines.data(); n = H.lines.size() * sizeof(LineCoef); break;
    case 5: src = H.xviews.data(); n = H.xviews.size() * sizeof(int); break;
    case 6: src = H.yviews.data(); n = H.yviews.size() * sizeof(int); break;
    case 7: src = H.ramp_taps.data(); n = H.ramp_taps.size() * sizeof(double); break;
    case 8: src = H.ramp_preweight.data(); n = H.ramp_preweight.size() * sizeof(double); break;
    case 10: src = H.adj_lk.data(); n = H.adj_lk.size() * sizeof(float); break;
    case 11: src = H.adj_wk.data(); n = H.adj_wk.size() * sizeof(float); break;
    case 9:
        ea[0] = H.ea_c[0]; ea[1] = H.ea_c[1]; ea[2] = H.ea_c[2]; ea[3] = H.ea_c[3]; ea[4] = H.ea_c[4];
        ea[5] = H.ea_c[5]; ea[6] = H.ea_T; ea[7] = (float)H.ea_poly;
        src = ea; n = sizeof(ea);
        break;
    default: return fail(CTLIB_ERR_INVALID_ARG, "unknown table");
    }
    if (dst) {
        if (*bytes < n) return fail(CTLIB_ERR_INVALID_ARG, "destination too small");
        if (n) std::memcpy(dst, src, n);
    }
    *bytes = n;
    return CTLIB_OK;
}

int ctlib_run(const ctlib_plan* p, int op, int dtype, const void* in, void* out, int64_t batch, void* ws,
              size_t ws_bytes, void* stream) {
    int rc = check_run_args(p, op, dtype, in, out, batch);
    if (rc) return rc;
    if (batch == 0) return CTLIB_OK;
    const WsLayout L = layout(p, snap(p), op, batch, dtype);
    rc = check_ws(L.total, ws, ws_bytes, "ctlib_workspace_bytes()");
    if (rc) return rc;
    return run_with(p, op, dtype, in, out, batch, L, static_cast<char*>(ws), static_cast<hipStream_t>(stream));
}

int ctlib_run_noalloc(const ctlib_plan* p, int op, int dtype, const void* in, void* out, int64_t batch,
                      void* stream) {
    int rc = check_run_args(p, op, dtype, in, out, batch);
    if (rc) return rc;
    if (batch == 0) return CTLIB_OK;
    // the automatic launch split and prebuilt line tables when they fit the scratch array, else
    // tables built per block, then fewer chunks
    Tuning T = snap(p);
    WsLayout L = layout(p, T, op, batch, dtype);
    if (L.total > kScratchBytes && L.n_ltab > 0) {
        T.fwd_lt = 1;
        L = layout(p, T, op, batch, dtype);
    }
    while (L.total > kScratchBytes && (L.fnch > 1 || L.nch > 1)) {
        T.fwd_nch = L.fnch > 1 ? L.fnch / 2 : 1;
        T.adj_nch = L.nch > 1 ? L.nch / 2 : 1;
        L = layout(p, T, op, batch, dtype);
    }
    if (L.total > kScratchBytes)
        return fail(CTLIB_ERR_WORKSPACE, "ctlib_run_noalloc: this op and batch need " + std::to_string(L.total) +
                                             " bytes of scratch, above its fixed " + std::to_string(kScratchBytes) +
                                             " (use ctlib_run with a workspace, or a smaller batch)");
    void* ptr = scratch_ptr(p->device);
    if (L.total > 0 && !ptr)
        return fail(CTLIB_ERR_HIP, "ctlib_run_noalloc: the scratch array's address on device " +
                                       std::to_string(p->device) + " is unknown (hipGetSymbolAddress failed)");
    NoTiming quiet;
    return run_with(p, op, dtype, in, out, batch, L, static_cast<char*>(ptr), static_cast<hipStream_t>(stream));
}

size_t ctlib_laplacian_workspace_bytes(int64_t n, int d, int k, int dtype) {
    if (n <= 0 || k <= 0 || d <= 0) return 0;
    return laplacian_ws_bytes(n, d, k, dtype == CTLIB_F64 ? CTLIB_F64 : CTLIB_F32);
}

int ctlib_laplacian(const void* x, int64_t n, int d, int k, int dtype, int64_t* cols, float* vals, void* ws,
                    size_t ws_bytes, void* stream) {
    if (dtype != CTLIB_F32 && dtype != CTLIB_F64) return fail(CTLIB_ERR_UNSUPPORTED_DTYPE, "dtype must be f32 or f64");
    if (n < 1 || n > 0x7fffffff || d < 1) return fail(CTLIB_ERR_INVALID_ARG, "need 1 <= points < 2^31, dimension >= 1");
    if (k < 1 || k > n) return fail(CTLIB_ERR_INVALID_ARG, "k must be in [1, number of points]");
    if (!x || !cols || !vals) return fail(CTLIB_ERR_INVALID_ARG, "NULL pointer");
    if (!laplacian_supported(d, k, dtype)) return fail(CTLIB_ERR_INVALID_ARG, "supported: k <= 4096");
    if (!ws || ws_bytes < ctlib_laplacian_workspace_bytes(n, d, k, dtype))
        return fail(CTLIB_ERR_WORKSPACE, "workspace smaller than ctlib_laplacian_workspace_bytes()");
    TraceRange range("ctlib:laplacian");
    hipError_t e = launch_laplacian(x, n, d, k, dtype, cols, vals, ws, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "laplacian launch");
    return CTLIB_OK;
}

int ctlib_timing_enable(int on) {
    std::lock_guard<std::mutex> lock(g_tmu);
    g_timing = on != 0;
    return CTLIB_OK;
}

int ctlib_timing_collect(ctlib_kernel_stats* out) {
    if (!out) return fail(CTLIB_ERR_INVALID_ARG, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    std::lock_guard<std::mutex> lock(g_tmu);
    int rc = CTLIB_OK;
    for (const TimedLaunch& t : g_rec) {
        float ms = 0.f;
        hipError_t e = hipEventSynchronize(t.b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t.a, t.b);
        if (e != hipSuccess) {
            rc = hip_fail(e, "timing events");
        } else {
            out->launches[t.cls] += 1;
            out->ms[t.cls] += ms;
            out->algo_bytes[t.cls] += t.bytes;
            out->flops[t.cls] += t.flops;
        }
        g_pool.push_back(t.a);
        g_pool.push_back(t.b);
    }
    g_rec.clear();
    return rc;
}

}  // extern "C"

namespace {

// Scratch of the residual step: [Y-line copy | partial sinograms] (projection's layout) and
// [Y-partial image] (projection_t's) overlapping, then the bin tables.
struct ResidLayout {
    WsLayout la, lb;
    size_t tables = 0, total = 0;
};

ResidLayout resid_layout(const ctlib_plan* p, const Tuning& T, int64_t batch, int dtype) {
    ResidLayout R;
    R.la = layout(p, T, CTLIB_OP_PRJ, batch, dtype);
    R.lb = layout(p, T, CTLIB_OP_PRJ_T, batch, dtype);
    R.tables = R.la.total > R.lb.total ? R.la.total : R.lb.total;
    R.total = R.tables + align_up((size_t)batch * p->dev.V * (p->dev.D + 2) * sizeof(double2));
    return R;
}

}  // namespace

extern "C" {

size_t ctlib_residual_workspace_bytes(const ctlib_plan* p, int64_t batch, int dtype) {
    if (!p || batch <= 0) return 0;
    return resid_layout(p, snap(p), batch, dtype).total;
}

int ctlib_run_residual(const ctlib_plan* p, int dtype, const void* x, const void* m, void* y, void* grad,
                       int64_t batch, void* ws, size_t ws_bytes, void* stream) {
    return ctlib_run_residual_loss(p, dtype, x, m, y, grad, nullptr, batch, ws, ws_bytes, stream);
}

int ctlib_run_residual_loss(const ctlib_plan* p, int dtype, const void* x, const void* m, void* y, void* grad,
                            double* row_sq, int64_t batch, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_run_args(p, CTLIB_OP_PRJ_T, dtype, x, grad, batch);
    if (rc) return rc;
    if (batch == 0) return CTLIB_OK;
    if (!m) return fail(CTLIB_ERR_INVALID_ARG, "measurement pointer is NULL");
    const ResidLayout R = resid_layout(p, snap(p), batch, dtype);
    rc = check_ws(R.total, ws, ws_bytes, "ctlib_residual_workspace_bytes()");
    if (rc) return rc;
    DeviceGuard guard(p->device);
    if (!guard.ok) return fail(CTLIB_ERR_HIP, "cannot select the plan's device");
    TraceRange range("ctlib:residual_step");
    const DevPlan& P = p->dev;
    const int B = (int)batch;
    const int dt = dtype == CTLIB_F64 ? DT_F64 : DT_F32;
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* w = static_cast<char*>(ws);
    const WsLayout& la = R.la;
    const WsLayout& lb = R.lb;
    double2* tabs = reinterpret_cast<double2*>(w + R.tables);
    void* yl = w + la.ylines;
    hipError_t e = hipSuccess;
    if (P.ny) {
        KernelTimer t(K_AUX, 0, 0, s);
        e = launch_to_ylines(x, yl, dt, B, P.H, P.W, s);
    }
    double2* lt = la.line_tables(w);
    if (e == hipSuccess && lt) e = run_line_tables(P, dt, x, P.ny ? yl : nullptr, lt, B, s);
    if (e == hipSuccess && la.fnch > 1) {
        // Batch too small to fill the chip unsplit: the forward is split into line chunks, and
        // one pass sums the partial sinograms (FP64, chunk order), writes y if wanted and forms the
        // residual tables and row sums.  (Forming them instead in the last chunk of every work
        // item, on completion counters, measured 0.17 ms slower at fan-ED 512^2 B=32.)
        {
            KernelTimer t(K_FWD, sweep_bytes(P, B, dt), 0, s);
            e = launch_fwd(P, dt, x, P.ny ? yl : nullptr, w + la.sp, B, false, P.post_prj, fwd_shape(P, false, la), s,
                           lt);
        }
        if (e == hipSuccess) {
            KernelTimer t(K_AUX, 0, 0, s);
            e = launch_bin_tables(P, dt, w + la.sp, m, P.post_prj, 1.0, tabs, B, s, row_sq, la.fnch,
                                  (size_t)B * P.V * P.D, y);
        }
    } else if (e == hipSuccess) {
        // forward fused with the residual tables (no separate table pass)
        KernelTimer t(K_FWD, sweep_bytes(P, B, dt), 0, s);
        e = launch_fwd_resid(P, dt, x, P.ny ? yl : nullptr, y, m, tabs, B, fwd_shape(P, false, la), s, row_sq, lt);
    }
    if (e == hipSuccess) e = run_to_image(p, dt, tabs, grad, B, false, 1.0, lb, w, s);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return CTLIB_OK;
}

}  // extern "C"
