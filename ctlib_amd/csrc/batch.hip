// Heterogeneous batch: slices of different image size, view count, detector
// count and scan type, run as one persistent fused forward + adjoint launch.
//
// Per slice s the work is  y_s = A_s x_s,  grad_s = A_s^T (A_s x_s - m_s),
// split into work items of five phases:
//   T  band of 64 image rows -> Y-axis line copy               (slices with Y-axis views)
//   L  16 image lines of one sweep axis -> their cumulative line tables
//   F  G=4 views of one sweep axis: projection over the line tables + residual bin tables
//   A  tile of lines of one sweep axis: projection_t over all views of the axis
//   C  band of 64 image rows: X part + transposed Y part       (slices with Y-axis views)
// The item list is ordered phase by phase (within a phase by the tuned order)
// and dequeued with one atomic counter by a fixed grid of resident blocks.  An
// item waits only for items of an earlier phase of its own slice, which were
// dequeued before it by running blocks, so the wait always ends: per-slice
// completion counters are released (agent scope) by the producer and acquired
// by the consumer.  The launch allocates nothing, synchronises nothing and
// records no memset (counters are reset by a kernel), so it can be captured in
// a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ctlib_abi.h"
#include "dd_device.h"
#include "kernels.h"

namespace ctlib {

using namespace dd;

namespace {

enum { IT_T = 0, IT_L = 1, IT_F = 2, IT_A = 3, IT_C = 4, NPHASE = 5 };
constexpr int BAND = 64;   // image rows per T / C item
constexpr int BG = 4;      // views per F item
constexpr int LL = 16;     // image lines per L item (4 per wave)

struct SliceDev {
    DevPlan P;
    const void* x;
    const void* m;
    void* y;
    void* grad;
    void* xT;
    double2* tabs;
    void* yt;
    double2* lt;  // line tables: X-axis lines then Y-axis lines, ltab_n<M>() entries each
    int fcls;   // forward variant (edge-slot class, see run_fwd)
    int acls;   // adjoint variant: line class * 4 + GEO (GEO_EAP for fan-EA with the polynomial atan)
    int ngx, ngy, ntx, nty, nband;
    int nlx, nly;  // L items of each axis
    int tileL;  // image lines per adjoint item
    int pad[3];
};

struct Item {
    int type, slice, mode, idx;
};


// Pixels per thread of the line tables: M = 2 (lines <= 512) or, in a batch with a slice beyond
// 512 x 512 (LARGE), M = 4 for every slice (lines <= 1024).
template <bool LARGE>
constexpr int batch_px() { return LARGE ? 4 : 2; }

// Line tables of one L item: lines idx*LL .. idx*LL + LL - 1 of one axis, four per wave.
template <typename T, bool LARGE>
__device__ __forceinline__ void run_lines(const SliceDev& sd, int mode, int idx) {
    constexpr int M = batch_px<LARGE>();
    const DevPlan& P = sd.P;
    const int nl = mode == MODE_Y ? P.W : P.H, n = mode == MODE_Y ? P.H : P.W;
    const T* src = (const T*)(mode == MODE_Y ? sd.xT : sd.x);
    double2* lt = sd.lt + (size_t)(mode == MODE_Y && P.nx ? P.H : 0) * ltab_n<M>();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < LL / 4; k++) {
        const int l = idx * LL + wave * (LL / 4) + k;
        if (l < nl) line_table_wave<T, M>(src + (size_t)l * n, n, lt + (size_t)l * ltab_n<M>(), lane);
    }
}

// Forward of one F item over the slice's line tables.  Edge slots per view: 12 (D < 768), 20
// (D < 1280) or 24 (D < 1536); 512, 1024 or 1536 bins: 8, 16 or 24 slots and the last edge packed.
template <typename T, bool LARGE>
__device__ __forceinline__ void run_fwd(const SliceDev& sd, int mode, int grp, unsigned char* smem) {
    constexpr int M = batch_px<LARGE>();
    T* y = (T*)sd.y;
    const T* m = (const T*)sd.m;
    const DevPlan& P = sd.P;
    const double2* lt = sd.lt + (size_t)(mode == MODE_Y && P.nx ? P.H : 0) * ltab_n<M>();
    switch (sd.fcls) {
    case 0: fwd_block_lt<T, 1, 12, M, true, false, 2>(P, lt, y, 0, mode, grp, P.post_prj, m, sd.tabs, smem, 0, -1, nullptr); break;
    case 2: fwd_block_lt<T, 1, 8, M, true, true, 2>(P, lt, y, 0, mode, grp, P.post_prj, m, sd.tabs, smem, 0, -1, nullptr); break;
    case 3: fwd_block_lt<T, 1, 16, M, true, true, 2>(P, lt, y, 0, mode, grp, P.post_prj, m, sd.tabs, smem, 0, -1, nullptr); break;
    case 4:
        if constexpr (LARGE) fwd_block_lt<T, 1, 24, M, true, false, 2>(P, lt, y, 0, mode, grp, P.post_prj, m, sd.tabs, smem, 0, -1, nullptr);
        break;
    case 5:
        if constexpr (LARGE) fwd_block_lt<T, 1, 24, M, true, true, 2>(P, lt, y, 0, mode, grp, P.post_prj, m, sd.tabs, smem, 0, -1, nullptr);
        break;
    default: fwd_block_lt<T, 1, 20, M, true, false, 2>(P, lt, y, 0, mode, grp, P.post_prj, m, sd.tabs, smem, 0, -1, nullptr); break;
    }
}

template <typename T, int LPW, int NPE, bool EXACT>
__device__ __forceinline__ void run_adj_geo(const SliceDev& sd, int mode, int tile, unsigned char* smem) {
    T* ox = (T*)sd.grad;
    T* oy = (T*)sd.yt;
    switch (sd.acls & 3) {
    case GEO_ED: adj_block<T, LPW, NPE, false, GEO_ED, EXACT>(sd.P, sd.tabs, ox, oy, 0, mode, tile, 1.0, smem); break;
    case GEO_EA: adj_block<T, LPW, NPE, false, GEO_EA, EXACT>(sd.P, sd.tabs, ox, oy, 0, mode, tile, 1.0, smem); break;
    case GEO_PARA:
        adj_block<T, LPW, NPE, false, GEO_PARA, EXACT>(sd.P, sd.tabs, ox, oy, 0, mode, tile, 1.0, smem);
        break;
    default: adj_block<T, LPW, NPE, false, GEO_EAP, EXACT>(sd.P, sd.tabs, ox, oy, 0, mode, tile, 1.0, smem); break;
    }
}

// Line classes: 0 lines <= 256 px, 1 lines <= 512 px; all lines exactly 2: 256, 3: 512, 4: 384 px
// (the last pixel edge evaluated packed); LARGE batches also 5: lines <= 1024 px, 6: exactly 1024 px.
template <typename T, bool LARGE>
__device__ __forceinline__ void run_adj(const SliceDev& sd, int mode, int tile, unsigned char* smem) {
    switch (sd.acls >> 2) {
    case 0: run_adj_geo<T, 8, 5, false>(sd, mode, tile, smem); break;
    case 2: run_adj_geo<T, 8, 5, true>(sd, mode, tile, smem); break;
    case 3: run_adj_geo<T, 4, 9, true>(sd, mode, tile, smem); break;
    case 4: run_adj_geo<T, 4, 7, true>(sd, mode, tile, smem); break;
    case 5:
        if constexpr (LARGE) run_adj_geo<T, 2, 17, false>(sd, mode, tile, smem);
        break;
    case 6:
        if constexpr (LARGE) run_adj_geo<T, 2, 17, true>(sd, mode, tile, smem);
        break;
    default: run_adj_geo<T, 4, 9, false>(sd, mode, tile, smem); break;
    }
}

template <typename T>
__device__ __forceinline__ void run_band(const SliceDev& sd, int type, int band, T (*tile)[33]) {
    const int H = sd.P.H, W = sd.P.W;
    const int r_end = min(H, (band + 1) * BAND);
    for (int r0 = band * BAND; r0 < r_end; r0 += 32) {
        for (int c0 = 0; c0 < W; c0 += 32) {
            if (type == IT_T)
                ylines_tile<T>((const T*)sd.x, (T*)sd.xT, H, W, r0, c0, tile);
            else
                combine_tile<T>((T*)sd.grad, (const T*)sd.yt, H, W, r0, c0, sd.P.nx > 0, 1, tile);
            __syncthreads();
        }
    }
}

// ctr layout: [0] dequeue counter, [1] wait-timeout flag (own 256-byte line), then per-slice done
// counters of T, L, F and A.
__global__ void k_batch_reset(unsigned int* ctr, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctr[i] = 0u;
}

// Register budget: 128 VGPRs (four waves per SIMD) for batches of slices within 512^2 and 1279
// bins - 7.22 ms for the 64-slice batch against 7.66 ms at 168 (three waves) - and 168 for the
// LARGE variant, whose 24-slot forward needs more.
template <typename T, bool LARGE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LARGE ? 3 : 4))) void k_batch(const SliceDev* __restrict__ slices, const Item* __restrict__ items,
                                               int nitems, unsigned int* __restrict__ ctr, int nslices) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_item;
    unsigned int* next = ctr;
    unsigned int* tdone = ctr + 64;
    unsigned int* ldone = tdone + nslices;
    unsigned int* fdone = ldone + nslices;
    unsigned int* adone = fdone + nslices;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(next, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const Item I = items[it];
        const SliceDev& sd = slices[I.slice];
        if (threadIdx.x == 0) {
            const unsigned int* c = nullptr;
            unsigned int need = 0;
            if (I.type == IT_L && I.mode == MODE_Y) {
                c = tdone + I.slice;
                need = sd.nband;
            } else if (I.type == IT_F) {
                c = ldone + I.slice;
                need = sd.nlx + sd.nly;
            } else if (I.type == IT_A) {
                c = fdone + I.slice;
                need = sd.ngx + sd.ngy;
            } else if (I.type == IT_C) {
                c = adone + I.slice;
                need = sd.ntx + sd.nty;
            }
            if (c) {
                // Bounded wait (about 2 s at most): every dependency is an item dequeued earlier by a
                // running block, so the count always arrives; the bound only turns a bookkeeping error
                // into a flagged wrong result instead of a hang.
                unsigned int spins = 0;
                while (__hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    __builtin_amdgcn_s_sleep(4);
                    if (++spins > (1u << 23)) {
                        __hip_atomic_store(ctr + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
        }
        __syncthreads();
        unsigned int* done = nullptr;
        switch (I.type) {
        case IT_T:
            run_band<T>(sd, IT_T, I.idx, reinterpret_cast<T(*)[33]>(smem));
            done = tdone + I.slice;
            break;
        case IT_L:
            run_lines<T, LARGE>(sd, I.mode, I.idx);
            done = ldone + I.slice;
            break;
        case IT_F:
            run_fwd<T, LARGE>(sd, I.mode, I.idx, smem);
            done = fdone + I.slice;
            break;
        case IT_A:
            run_adj<T, LARGE>(sd, I.mode, I.idx, smem);
            done = adone + I.slice;
            break;
        default:
            run_band<T>(sd, IT_C, I.idx, reinterpret_cast<T(*)[33]>(smem));
            break;
        }
        // The barrier completes every wave's stores (write-through L1 -> L2) before thread 0's
        // agent-scope release, which makes them visible to the acquiring block on any XCD.
        __syncthreads();
        if (threadIdx.x == 0 && done) __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace ctlib

using namespace ctlib;

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct ctlib_batch {
    int device = 0;
    int dtype = CTLIB_F32;
    int nslices = 0;
    std::vector<ctlib_plan*> plans;
    std::vector<SliceDev> slices;  // host copy (pointers filled by bind)
    std::vector<Item> items;
    std::vector<double> item_cost;
    SliceDev* d_slices = nullptr;
    Item* d_items = nullptr;
    unsigned int* d_ctr = nullptr;
    int nctr = 0;
    size_t ws_bytes = 0;
    std::vector<size_t> ws_off;    // per slice: xT, tabs, yt, line-table offsets
    int blocks_per_cu = 2;
    int order = 0;
    int num_cu = 256;
    size_t lds = 0;
    bool large = false;  // a slice beyond 512 x 512 or 1279 bins: the LARGE kernel variant
    bool bound = false;
};

namespace {

thread_local std::string b_err;

int bfail(int code, const std::string& why) {
    b_err = why;
    return code;
}

size_t esize(int dtype) { return dtype == CTLIB_F64 ? 8 : 4; }

void build_items(ctlib_batch* b) {
    b->items.clear();
    b->item_cost.clear();
    std::vector<std::pair<double, Item>> ph[NPHASE];
    for (int s = 0; s < b->nslices; s++) {
        const SliceDev& sd = b->slices[s];
        const DevPlan& P = sd.P;
        if (P.ny)
            for (int i = 0; i < sd.nband; i++) ph[IT_T].push_back({(double)BAND * P.W, Item{IT_T, s, 0, i}});
        for (int i = 0; i < sd.nlx; i++) ph[IT_L].push_back({(double)LL * P.W, Item{IT_L, s, MODE_X, i}});
        for (int i = 0; i < sd.nly; i++) ph[IT_L].push_back({(double)LL * P.H, Item{IT_L, s, MODE_Y, i}});
        for (int g = 0; g < sd.ngx; g++)
            ph[IT_F].push_back({(double)P.H * (P.D + 1) * BG, Item{IT_F, s, MODE_X, g}});
        for (int g = 0; g < sd.ngy; g++)
            ph[IT_F].push_back({(double)P.W * (P.D + 1) * BG, Item{IT_F, s, MODE_Y, g}});
        const int L = sd.tileL;
        for (int t = 0; t < sd.ntx; t++) ph[IT_A].push_back({(double)L * (P.W + 1) * P.nx, Item{IT_A, s, MODE_X, t}});
        for (int t = 0; t < sd.nty; t++) ph[IT_A].push_back({(double)L * (P.H + 1) * P.ny, Item{IT_A, s, MODE_Y, t}});
        if (P.ny)
            for (int i = 0; i < sd.nband; i++) ph[IT_C].push_back({(double)BAND * P.W, Item{IT_C, s, 0, i}});
    }
    for (int p = 0; p < NPHASE; p++) {
        if (b->order == 0) {  // longest item first within the phase
            std::stable_sort(ph[p].begin(), ph[p].end(),
                             [](const std::pair<double, Item>& a, const std::pair<double, Item>& c) {
                                 return a.first > c.first;
                             });
        }
        for (auto& pi : ph[p]) {
            b->items.push_back(pi.second);
            b->item_cost.push_back(pi.first);
        }
    }
}

}  // namespace

extern "C" {

const char* ctlib_batch_last_error(void) { return b_err.c_str(); }

int ctlib_batch_create(const ctlib_geom* geoms, int nslices, int device, int dtype, ctlib_batch** out) {
    if (!geoms || !out || nslices < 1) return bfail(CTLIB_ERR_INVALID_ARG, "bad arguments");
    if (dtype != CTLIB_F32 && dtype != CTLIB_F64) return bfail(CTLIB_ERR_UNSUPPORTED_DTYPE, "dtype must be f32/f64");
    *out = nullptr;
    ctlib_batch* b = new ctlib_batch();
    b->device = device;
    b->dtype = dtype;
    b->nslices = nslices;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    auto cleanup = [&](int code, const std::string& why) {
        for (ctlib_plan* p : b->plans) ctlib_plan_destroy(p);
        delete b;
        if (prev >= 0) (void)hipSetDevice(prev);
        return bfail(code, why);
    };
    size_t off = 0;
    std::vector<size_t> lines_total;
    for (int s = 0; s < nslices; s++) {
        const ctlib_geom& g = geoms[s];
        if (g.width > 1024 || g.height > 1024 || g.dets > 1536)
            return cleanup(CTLIB_ERR_GEOMETRY, "batched slices support width, height <= 1024 and dets <= 1536");
        if (g.width > 512 || g.height > 512 || g.dets > 1279) b->large = true;
        ctlib_plan* p = nullptr;
        int rc = ctlib_plan_create(&g, device, &p);
        if (rc) return cleanup(rc, std::string("slice ") + std::to_string(s) + ": " + ctlib_last_error());
        b->plans.push_back(p);
        SliceDev sd{};
        sd.P = plan_dev(p);
        const DevPlan& P = sd.P;
        const int ne = (P.D + 1 + 63) / 64;
        const int nmax = std::max(P.nx ? P.W : 0, P.ny ? P.H : 0);
        // forward class (run_fwd; 4 and 5 occur only in LARGE batches)
        sd.fcls = P.D == 512 ? 2 : P.D == 1024 ? 3 : P.D == 1536 ? 5 : ne <= 12 ? 0 : ne <= 20 ? 1 : 4;
        const int npe = (nmax + 1 + 63) / 64;
        // every line (rows of X views, columns of Y views) exactly n pixels
        auto all_lines = [&](int n) { return (P.nx == 0 || P.W == n) && (P.ny == 0 || P.H == n); };
        const int lcls = all_lines(256)    ? 2
                         : all_lines(512)  ? 3
                         : all_lines(384)  ? 4
                         : all_lines(1024) ? 6
                         : npe <= 5        ? 0
                         : npe <= 9        ? 1
                                           : 5;
        sd.acls = lcls * 4 + (P.scan_type == GEO_EA && P.ea_poly ? GEO_EAP : P.scan_type);
        const int L = lcls == 0 || lcls == 2 ? 32 : lcls >= 5 ? 8 : 16;
        sd.tileL = L;
        sd.ngx = (P.nx + BG - 1) / BG;
        sd.ngy = (P.ny + BG - 1) / BG;
        sd.ntx = P.nx ? (P.H + L - 1) / L : 0;
        sd.nty = P.ny ? (P.W + L - 1) / L : 0;
        sd.nband = (P.H + BAND - 1) / BAND;
        sd.nlx = P.nx ? (P.H + LL - 1) / LL : 0;
        sd.nly = P.ny ? (P.W + LL - 1) / LL : 0;
        b->slices.push_back(sd);
        const size_t es = esize(dtype);
        b->ws_off.push_back(off);
        off += align_up((size_t)P.H * P.W * es);                   // xT
        b->ws_off.push_back(off);
        off += align_up((size_t)P.V * (P.D + 2) * sizeof(double2));  // tabs
        b->ws_off.push_back(off);
        off += align_up((size_t)P.H * P.W * es);                   // yt
        b->ws_off.push_back(0);  // line tables: placed once every slice's size class is known
        lines_total.push_back((size_t)(P.nx ? P.H : 0) + (P.ny ? P.W : 0));
        b->lds = std::max(b->lds, adj_block_lds(P.D, L));
    }
    // line tables last (their entry count per line depends on the batch's pixels-per-thread class)
    {
        const size_t tabn = b->large ? ltab_n<batch_px<true>()>() : ltab_n<batch_px<false>()>();
        for (int s = 0; s < nslices; s++) {
            b->ws_off[4 * s + 3] = off;
            off += align_up(lines_total[s] * tabn * sizeof(double2));
        }
        b->lds = std::max(b->lds, b->large ? fwd_lt_lds<batch_px<true>()>(BG, 2) : fwd_lt_lds<batch_px<false>()>(BG, 2));
    }
    b->lds = std::max(b->lds, (size_t)32 * 33 * 8);
    b->ws_bytes = off;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->num_cu = prop.multiProcessorCount;
    build_items(b);
    b->nctr = 64 + 4 * nslices;
    if (hipMalloc(&b->d_slices, sizeof(SliceDev) * nslices) != hipSuccess ||
        hipMalloc(&b->d_items, sizeof(Item) * b->items.size()) != hipSuccess ||
        hipMalloc(&b->d_ctr, sizeof(unsigned int) * b->nctr) != hipSuccess)
        return cleanup(CTLIB_ERR_HIP, "hipMalloc(batch)");
    if (hipMemcpy(b->d_items, b->items.data(), sizeof(Item) * b->items.size(), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(CTLIB_ERR_HIP, "hipMemcpy(batch items)");
    if (prev >= 0) (void)hipSetDevice(prev);
    *out = b;
    return CTLIB_OK;
}

void ctlib_batch_destroy(ctlib_batch* b) {
    if (!b) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    if (b->d_slices) (void)hipFree(b->d_slices);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_ctr) (void)hipFree(b->d_ctr);
    for (ctlib_plan* p : b->plans) ctlib_plan_destroy(p);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete b;
}

size_t ctlib_batch_workspace_bytes(const ctlib_batch* b) { return b ? b->ws_bytes : 0; }

int ctlib_batch_set_schedule(ctlib_batch* b, int blocks_per_cu, int order) {
    if (!b || blocks_per_cu < 1 || blocks_per_cu > 8 || order < 0 || order > 1)
        return bfail(CTLIB_ERR_INVALID_ARG, "blocks_per_cu in 1..8, order 0 (longest first) or 1 (slice order)");
    b->blocks_per_cu = blocks_per_cu;
    if (order != b->order) {
        b->order = order;
        build_items(b);
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(b->device);
        hipError_t e = hipMemcpy(b->d_items, b->items.data(), sizeof(Item) * b->items.size(), hipMemcpyHostToDevice);
        if (prev >= 0) (void)hipSetDevice(prev);
        if (e != hipSuccess) return bfail(CTLIB_ERR_HIP, "hipMemcpy(batch items)");
    }
    return CTLIB_OK;
}

int ctlib_batch_bind(ctlib_batch* b, const void* const* x, const void* const* m, void* const* y, void* const* grad,
                     void* ws, size_t ws_bytes) {
    if (!b || !x || !m || !grad) return bfail(CTLIB_ERR_INVALID_ARG, "NULL argument");
    if (!ws || ws_bytes < b->ws_bytes) return bfail(CTLIB_ERR_WORKSPACE, "workspace smaller than required");
    char* w = static_cast<char*>(ws);
    for (int s = 0; s < b->nslices; s++) {
        if (!x[s] || !m[s] || !grad[s]) return bfail(CTLIB_ERR_INVALID_ARG, "NULL slice pointer");
        SliceDev& sd = b->slices[s];
        sd.x = x[s];
        sd.m = m[s];
        sd.y = y ? y[s] : nullptr;
        sd.grad = grad[s];
        sd.xT = w + b->ws_off[4 * s];
        sd.tabs = reinterpret_cast<double2*>(w + b->ws_off[4 * s + 1]);
        sd.yt = w + b->ws_off[4 * s + 2];
        sd.lt = reinterpret_cast<double2*>(w + b->ws_off[4 * s + 3]);
    }
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    hipError_t e = hipMemcpy(b->d_slices, b->slices.data(), sizeof(SliceDev) * b->nslices, hipMemcpyHostToDevice);
    if (prev >= 0) (void)hipSetDevice(prev);
    if (e != hipSuccess) return bfail(CTLIB_ERR_HIP, "hipMemcpy(batch slices)");
    b->bound = true;
    return CTLIB_OK;
}

int ctlib_batch_run(const ctlib_batch* b, void* stream) {
    if (!b || !b->bound) return bfail(CTLIB_ERR_INVALID_ARG, "batch not bound");
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    roctxRangePushA("ctlib:batch_residual_step");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(256), 0, s, b->d_ctr, b->nctr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        auto kernel = [&](auto kern) {
            if (b->lds > 64 * 1024)
                e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds);
            if (e != hipSuccess) return;
            hipLaunchKernelGGL(kern, dim3(b->num_cu * b->blocks_per_cu), dim3(256), b->lds, s, b->d_slices,
                               b->d_items, (int)b->items.size(), b->d_ctr, b->nslices);
            e = hipGetLastError();
        };
        if (b->dtype == CTLIB_F64)
            b->large ? kernel(k_batch<double, true>) : kernel(k_batch<double, false>);
        else
            b->large ? kernel(k_batch<float, true>) : kernel(k_batch<float, false>);
    }
    roctxRangePop();
    if (prev >= 0) (void)hipSetDevice(prev);
    if (e != hipSuccess) return bfail(CTLIB_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
    return CTLIB_OK;
}

int ctlib_batch_check(const ctlib_batch* b, void* stream) {
    if (!b || !b->d_ctr) return bfail(CTLIB_ERR_INVALID_ARG, "NULL batch");
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    unsigned int flag = 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(&flag, b->d_ctr + 1, sizeof(flag), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (prev >= 0) (void)hipSetDevice(prev);
    if (e != hipSuccess) return bfail(CTLIB_ERR_HIP, std::string("batch check: ") + hipGetErrorString(e));
    if (flag) return bfail(CTLIB_ERR_HIP, "a batch item timed out waiting for its dependencies");
    return CTLIB_OK;
}

int ctlib_batch_info(const ctlib_batch* b, int* nitems, int* blocks_per_cu, int* order) {
    if (!b) return bfail(CTLIB_ERR_INVALID_ARG, "NULL batch");
    if (nitems) *nitems = (int)b->items.size();
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
    if (order) *order = b->order;
    return CTLIB_OK;
}

}  // extern "C"
