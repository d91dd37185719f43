// PyTorch-ROCm extension `ctlib_amd._C`: the reference's Python surface
// (ctlib.cpp:106-114) on top of the C ABI of libctlib_amd.so.
//
// Each op keeps the reference's signature f(Tensor, Tensor options) -> Tensor,
// tensor layout (B x 1 x H x W images, B x 1 x V x D sinograms, channel 0 read,
// output channel 1) and dtype set (float32, float64).  Differences:
//   * options may live on the host (no device sync) or on the device (one
//     D2H copy), in either float dtype; the reference needs device options of
//     the data dtype (fan_ed_kernel.cu:1188-1191);
//   * shapes are checked against the options (the reference reads out of
//     bounds on a mismatch) and bad scan_types raise instead of exit(0);
//   * kernels run on the current HIP stream under a device guard (the
//     reference uses the legacy default stream and leaves cudaSetDevice set).
#include <torch/extension.h>

#include <c10/hip/HIPGuard.h>
#include <c10/hip/HIPStream.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ctlib_abi.h"

namespace {

struct PlanDeleter {
    void operator()(ctlib_plan* p) const { ctlib_plan_destroy(p); }
};
using PlanPtr = std::shared_ptr<ctlib_plan>;

struct Cached {
    PlanPtr plan;
    ctlib_geom geom;
};

// Plan cache: at most kMaxPlans geometries per process, least recently used evicted first.
// A plan used while its stream is being captured (torch.cuda.graph, torch.compile's graph
// trees) has its device tables' addresses baked into the captured kernels, so it is pinned:
// kept alive until the process ends, whatever the cache later evicts or clears.
struct Slot {
    Cached c;
    uint64_t last_use = 0;
};

std::mutex g_mu;
std::map<std::string, Slot> g_cache;
std::vector<PlanPtr> g_pinned;
uint64_t g_tick = 0;
int64_t g_created = 0;
constexpr size_t kMaxPlans = 64;

[[noreturn]] void raise_status(int rc, const char* what) {
    std::string msg = std::string(what) + ": " + ctlib_status_string(rc) + ": " + ctlib_last_error();
    if (rc == CTLIB_ERR_INVALID_ARG || rc == CTLIB_ERR_BAD_SCAN_TYPE || rc == CTLIB_ERR_SHAPE ||
        rc == CTLIB_ERR_GEOMETRY)
        throw pybind11::value_error(msg);
    if (rc == CTLIB_ERR_UNSUPPORTED_DTYPE) throw pybind11::type_error(msg);
    throw std::runtime_error(msg);
}

std::vector<double> options_values(const torch::Tensor& options) {
    TORCH_CHECK(options.dim() == 1, "options must be a 1-D tensor");
    TORCH_CHECK(options.scalar_type() == torch::kFloat32 || options.scalar_type() == torch::kFloat64,
                "options must be float32 or float64");
    torch::Tensor h = options.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    const double* p = h.data_ptr<double>();
    return std::vector<double>(p, p + h.numel());
}

Cached get_plan(const torch::Tensor& options, int device) {
    std::vector<double> vals = options_values(options);
    ctlib_geom g{};
    int rc = ctlib_parse_options(vals.data(), (int)vals.size(), &g);
    if (rc) raise_status(rc, "options");
    std::string key(reinterpret_cast<const char*>(&g), sizeof(g));
    key.append(reinterpret_cast<const char*>(&device), sizeof(device));
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
        it->second.last_use = ++g_tick;
        return it->second.c;
    }
    ctlib_plan* raw = nullptr;
    rc = ctlib_plan_create(&g, device, &raw);
    if (rc) raise_status(rc, "plan");
    g_created++;
    if (g_cache.size() >= kMaxPlans) {
        auto lru = g_cache.begin();
        for (auto j = g_cache.begin(); j != g_cache.end(); ++j)
            if (j->second.last_use < lru->second.last_use) lru = j;
        g_cache.erase(lru);  // a pinned plan stays alive through g_pinned
    }
    Slot slot{Cached{PlanPtr(raw, PlanDeleter()), g}, ++g_tick};
    g_cache.emplace(key, slot);
    return slot.c;
}

// Pins `plan` for the life of the process when `s` is capturing a graph.
void pin_if_capturing(const PlanPtr& plan, hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess || st == hipStreamCaptureStatusNone) return;
    std::lock_guard<std::mutex> lock(g_mu);
    for (const PlanPtr& q : g_pinned)
        if (q.get() == plan.get()) return;
    g_pinned.push_back(plan);
}

int dtype_code(const torch::Tensor& t) {
    if (t.scalar_type() == torch::kFloat32) return CTLIB_F32;
    if (t.scalar_type() == torch::kFloat64) return CTLIB_F64;
    throw pybind11::type_error(std::string("ctlib ops support float32 and float64 tensors, got ") +
                               std::string(c10::toString(t.scalar_type())));
}

// CHECK_INPUT of ctlib.cpp:7-9, plus the shape checks the reference omits.
torch::Tensor check_input(const torch::Tensor& x, const char* name) {
    TORCH_CHECK(x.is_cuda(), name, " must be a GPU (HIP) tensor");
    TORCH_CHECK(x.is_contiguous(), name, " must be contiguous");
    TORCH_CHECK(x.dim() == 4, name, " must be 4-D (B x 1 x rows x cols), got ", x.dim(), "-D");
    if (x.size(1) != 1) return x.select(1, 0).unsqueeze(1).contiguous();  // reference reads channel 0
    return x;
}

enum Kind { IMG_TO_SINO, SINO_TO_IMG, SINO_TO_VIEWS };

torch::Tensor run_op(const torch::Tensor& in_, const torch::Tensor& options, int op, Kind kind, const char* name) {
    torch::Tensor in = check_input(in_, name);
    const int dt = dtype_code(in);
    const int device = in.get_device();
    c10::hip::HIPGuard guard((c10::DeviceIndex)device);
    Cached c = get_plan(options, device);
    const ctlib_geom& g = c.geom;
    const int64_t B = in.size(0);
    if (kind == IMG_TO_SINO) {
        if (in.size(2) != g.height || in.size(3) != g.width)
            throw pybind11::value_error(std::string(name) + " is " + std::to_string(in.size(2)) + "x" +
                                        std::to_string(in.size(3)) + " but options give height x width = " +
                                        std::to_string(g.height) + "x" + std::to_string(g.width));
    } else {
        if (in.size(2) != g.views || in.size(3) != g.dets)
            throw pybind11::value_error(std::string(name) + " is " + std::to_string(in.size(2)) + "x" +
                                        std::to_string(in.size(3)) + " but options give views x dets = " +
                                        std::to_string(g.views) + "x" + std::to_string(g.dets));
    }
    torch::Tensor out = kind == IMG_TO_SINO   ? torch::empty({B, 1, g.views, g.dets}, in.options())
                        : kind == SINO_TO_IMG ? torch::empty({B, 1, g.height, g.width}, in.options())
                                              : torch::empty({B, g.views, g.height, g.width}, in.options());
    if (B == 0) return out;
    const size_t wsb = ctlib_workspace_bytes(c.plan.get(), op, B, dt);
    torch::Tensor ws;
    if (wsb) ws = torch::empty({(int64_t)wsb}, in.options().dtype(torch::kUInt8));
    hipStream_t s = c10::hip::getCurrentHIPStream((c10::DeviceIndex)device).stream();
    pin_if_capturing(c.plan, s);
    int rc = ctlib_run(c.plan.get(), op, dt, in.data_ptr(), out.data_ptr(), B, wsb ? ws.data_ptr() : nullptr, wsb,
                       (void*)s);
    if (rc) raise_status(rc, name);
    return out;
}

// ctlib.cpp:56-69
torch::Tensor projection(torch::Tensor image, torch::Tensor options) {
    return run_op(image, options, CTLIB_OP_PRJ, IMG_TO_SINO, "image");
}
// ctlib.cpp:71-84
torch::Tensor projection_t(torch::Tensor projection, torch::Tensor options) {
    return run_op(projection, options, CTLIB_OP_PRJ_T, SINO_TO_IMG, "projection");
}
// ctlib.cpp:30-43
torch::Tensor backprojection(torch::Tensor projection, torch::Tensor options) {
    return run_op(projection, options, CTLIB_OP_BPRJ, SINO_TO_IMG, "projection");
}
// ctlib.cpp:15-28
torch::Tensor backprojection_t(torch::Tensor image, torch::Tensor options) {
    return run_op(image, options, CTLIB_OP_BPRJ_T, IMG_TO_SINO, "image");
}
// ctlib.cpp:86-99
torch::Tensor fbp(torch::Tensor projection, torch::Tensor options) {
    return run_op(projection, options, CTLIB_OP_FBP, SINO_TO_IMG, "projection");
}
// ctlib.cpp:45-54: per-view backprojection, output [B, V, H, W].  The reference
// implements it for scan_type 0 only and exit(0)s otherwise; that case raises
// here unless `any_scan_type` asks for the fan-EA / parallel extension.
torch::Tensor backprojection_sv(torch::Tensor projection, torch::Tensor options, bool any_scan_type) {
    if (!any_scan_type) {
        std::vector<double> vals = options_values(options);
        ctlib_geom g{};
        int rc = ctlib_parse_options(vals.data(), (int)vals.size(), &g);
        if (rc) raise_status(rc, "options");
        if (g.scan_type != 0)
            throw pybind11::value_error("backprojection_sv supports scan_type 0 (fan beam, equal distance) only, "
                                        "as the reference does");
    }
    return run_op(projection, options, CTLIB_OP_BPRJ_SV, SINO_TO_VIEWS, "projection");
}
// ctlib.cpp:101-104, laplacian_cuda_kernel.cu:46-128: kNN graph weights as a sparse COO
// (N, N) float32 tensor, row i holding its k nearest points (itself included).
torch::Tensor laplacian(torch::Tensor input, int64_t k) {
    TORCH_CHECK(input.is_cuda(), "input must be a GPU (HIP) tensor");
    TORCH_CHECK(input.is_contiguous(), "input must be contiguous");
    TORCH_CHECK(input.dim() == 2, "input must be 2-D (points x dimension), got ", input.dim(), "-D");
    const int dt = dtype_code(input);
    const int64_t n = input.size(0);
    const int d = (int)input.size(1);
    const int device = input.get_device();
    c10::hip::HIPGuard guard((c10::DeviceIndex)device);
    if (k < 1 || k > n) throw pybind11::value_error("k must be in [1, number of points]");
    torch::Tensor cols = torch::empty({n, k}, input.options().dtype(torch::kInt64));
    torch::Tensor vals = torch::empty({n, k}, input.options().dtype(torch::kFloat32));
    const size_t wsb = ctlib_laplacian_workspace_bytes(n, d, (int)k, dt);
    torch::Tensor ws = torch::empty({(int64_t)wsb}, input.options().dtype(torch::kUInt8));
    hipStream_t s = c10::hip::getCurrentHIPStream((c10::DeviceIndex)device).stream();
    int rc = ctlib_laplacian(input.data_ptr(), n, d, (int)k, dt, cols.data_ptr<int64_t>(), vals.data_ptr<float>(),
                             ws.data_ptr(), wsb, (void*)s);
    if (rc) raise_status(rc, "laplacian");
    torch::Tensor rows = torch::arange(n, cols.options()).unsqueeze(1).expand({n, k});
    torch::Tensor idx = torch::stack({rows.reshape({-1}), cols.reshape({-1})});
    return torch::sparse_coo_tensor(idx, vals.reshape({-1}), {n, n}, vals.options().layout(torch::kSparse));
}

// Exact adjoints used by autograd (no reference counterpart).
torch::Tensor backprojection_t_true(torch::Tensor image, torch::Tensor options) {
    return run_op(image, options, CTLIB_OP_BPRJ_T_TRUE, IMG_TO_SINO, "image");
}
torch::Tensor fbp_t(torch::Tensor image, torch::Tensor options) {
    return run_op(image, options, CTLIB_OP_FBP_T, IMG_TO_SINO, "image");
}

// One least-squares step: returns (A x, A^T (A x - m)); with_loss: (row sums of (A x - m)^2 as
// [B, V] f64, A^T (A x - m)) instead, without materialising A x.
std::vector<torch::Tensor> residual_common(torch::Tensor image, torch::Tensor measured, torch::Tensor options,
                                           bool with_loss) {
    torch::Tensor x = check_input(image, "image");
    torch::Tensor m = check_input(measured, "measured");
    const int dt = dtype_code(x);
    TORCH_CHECK(m.scalar_type() == x.scalar_type(), "measured must have the image dtype");
    TORCH_CHECK(m.get_device() == x.get_device(), "image and measured must be on the same device");
    const int device = x.get_device();
    c10::hip::HIPGuard guard((c10::DeviceIndex)device);
    Cached c = get_plan(options, device);
    const ctlib_geom& g = c.geom;
    const int64_t B = x.size(0);
    if (x.size(2) != g.height || x.size(3) != g.width)
        throw pybind11::value_error("image shape does not match the options");
    if (m.size(0) != B || m.size(2) != g.views || m.size(3) != g.dets)
        throw pybind11::value_error("measured sinogram shape does not match the options / batch");
    torch::Tensor first = with_loss ? torch::empty({B, g.views}, x.options().dtype(torch::kFloat64))
                                    : torch::empty({B, 1, g.views, g.dets}, x.options());
    torch::Tensor grad = torch::empty({B, 1, g.height, g.width}, x.options());
    if (B == 0) return {first, grad};
    const size_t wsb = ctlib_residual_workspace_bytes(c.plan.get(), B, dt);
    torch::Tensor ws = torch::empty({(int64_t)wsb}, x.options().dtype(torch::kUInt8));
    hipStream_t s = c10::hip::getCurrentHIPStream((c10::DeviceIndex)device).stream();
    pin_if_capturing(c.plan, s);
    int rc = ctlib_run_residual_loss(c.plan.get(), dt, x.data_ptr(), m.data_ptr(),
                                     with_loss ? nullptr : first.data_ptr(), grad.data_ptr(),
                                     with_loss ? first.data_ptr<double>() : nullptr, B, ws.data_ptr(), wsb, (void*)s);
    if (rc) raise_status(rc, "residual_step");
    return {first, grad};
}

std::vector<torch::Tensor> residual_step(torch::Tensor image, torch::Tensor measured, torch::Tensor options) {
    return residual_common(image, measured, options, false);
}

std::vector<torch::Tensor> residual_loss_step(torch::Tensor image, torch::Tensor measured, torch::Tensor options) {
    return residual_common(image, measured, options, true);
}

// Geometry of an options vector as the library parses it (for host-side code and tests).
pybind11::dict parse_options(torch::Tensor options) {
    std::vector<double> vals = options_values(options);
    ctlib_geom g{};
    int rc = ctlib_parse_options(vals.data(), (int)vals.size(), &g);
    if (rc) raise_status(rc, "options");
    pybind11::dict d;
    d["views"] = g.views;
    d["dets"] = g.dets;
    d["width"] = g.width;
    d["height"] = g.height;
    d["scan_type"] = g.scan_type;
    d["dImg"] = g.dImg;
    d["dDet"] = g.dDet;
    d["Ang0"] = g.Ang0;
    d["dAng"] = g.dAng;
    d["s2r"] = g.s2r;
    d["d2r"] = g.d2r;
    d["binshift"] = g.binshift;
    return d;
}

int64_t workspace_bytes(torch::Tensor options, int64_t op, int64_t batch, bool f64, int64_t device) {
    Cached c = get_plan(options, (int)device);
    return (int64_t)ctlib_workspace_bytes(c.plan.get(), (int)op, batch, f64 ? CTLIB_F64 : CTLIB_F32);
}

void set_plan_tuning(torch::Tensor options, int64_t device, int64_t fwd_views_per_block) {
    Cached c = get_plan(options, (int)device);
    int rc = ctlib_plan_set_tuning(c.plan.get(), (int)fwd_views_per_block);
    if (rc) raise_status(rc, "set_plan_tuning");
}

void set_plan_chunks(torch::Tensor options, int64_t device, int64_t line_chunks, int64_t view_chunks) {
    Cached c = get_plan(options, (int)device);
    int rc = ctlib_plan_set_chunks(c.plan.get(), (int)line_chunks, (int)view_chunks);
    if (rc) raise_status(rc, "set_plan_chunks");
}

void set_plan_line_tables(torch::Tensor options, int64_t device, int64_t per_block) {
    Cached c = get_plan(options, (int)device);
    int rc = ctlib_plan_set_line_tables(c.plan.get(), (int)per_block);
    if (rc) raise_status(rc, "set_plan_line_tables");
}

void timing_enable(bool on) { ctlib_timing_enable(on ? 1 : 0); }

// Per kernel class: launches, total milliseconds, algorithmic bytes and flops of the
// launches recorded since the last collect (classes: fwd, adj, ramp, aux, sv).
pybind11::dict timing_collect() {
    ctlib_kernel_stats st{};
    int rc = ctlib_timing_collect(&st);
    if (rc) raise_status(rc, "timing_collect");
    static const char* names[5] = {"fwd", "adj", "ramp", "aux", "sv"};
    pybind11::dict d;
    for (int i = 0; i < 5; i++) {
        pybind11::dict e;
        e["launches"] = st.launches[i];
        e["ms"] = st.ms[i];
        e["algo_bytes"] = st.algo_bytes[i];
        e["flops"] = st.flops[i];
        d[names[i]] = e;
    }
    return d;
}

// Heterogeneous batch (see ctlib_batch_* in the C ABI).
class Batch {
  public:
    Batch(std::vector<torch::Tensor> options, int64_t device, bool f64) : device_((int)device) {
        std::vector<ctlib_geom> geoms;
        for (auto& o : options) {
            std::vector<double> vals = options_values(o);
            ctlib_geom g{};
            int rc = ctlib_parse_options(vals.data(), (int)vals.size(), &g);
            if (rc) raise_status(rc, "options");
            geoms.push_back(g);
        }
        geoms_ = geoms;
        c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
        int rc = ctlib_batch_create(geoms.data(), (int)geoms.size(), device_, f64 ? CTLIB_F64 : CTLIB_F32, &b_);
        if (rc) throw std::runtime_error(std::string("ctlib_batch_create: ") + ctlib_batch_last_error());
    }
    ~Batch() { ctlib_batch_destroy(b_); }
    int64_t workspace_bytes() const { return (int64_t)ctlib_batch_workspace_bytes(b_); }
    void set_schedule(int64_t blocks_per_cu, int64_t order) {
        int rc = ctlib_batch_set_schedule(b_, (int)blocks_per_cu, (int)order);
        if (rc) throw pybind11::value_error(ctlib_batch_last_error());
    }
    void bind(std::vector<torch::Tensor> x, std::vector<torch::Tensor> m, std::vector<torch::Tensor> y,
              std::vector<torch::Tensor> grad, torch::Tensor ws) {
        const size_t n = geoms_.size();
        if (x.size() != n || m.size() != n || grad.size() != n || (!y.empty() && y.size() != n))
            throw pybind11::value_error("bind needs one tensor per slice");
        std::vector<const void*> px(n), pm(n);
        std::vector<void*> py(n), pg(n);
        for (size_t s = 0; s < n; s++) {
            const ctlib_geom& g = geoms_[s];
            auto chk = [&](const torch::Tensor& t, int64_t r, int64_t c, const char* what) {
                TORCH_CHECK(t.is_cuda() && t.is_contiguous() && t.get_device() == device_, what,
                            " must be a contiguous tensor on the batch's device");
                TORCH_CHECK(t.numel() == r * c, what, " of slice ", s, " must hold ", r, " x ", c, " elements");
            };
            chk(x[s], g.height, g.width, "x");
            chk(m[s], g.views, g.dets, "m");
            chk(grad[s], g.height, g.width, "grad");
            px[s] = x[s].data_ptr();
            pm[s] = m[s].data_ptr();
            pg[s] = grad[s].data_ptr();
            if (!y.empty()) {
                chk(y[s], g.views, g.dets, "y");
                py[s] = y[s].data_ptr();
            }
        }
        TORCH_CHECK(ws.is_cuda() && ws.get_device() == device_, "workspace must be on the batch's device");
        c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
        int rc = ctlib_batch_bind(b_, px.data(), pm.data(), y.empty() ? nullptr : py.data(), pg.data(), ws.data_ptr(),
                                  (size_t)ws.numel() * ws.element_size());
        if (rc) throw std::runtime_error(std::string("ctlib_batch_bind: ") + ctlib_batch_last_error());
    }
    void run() {
        c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
        hipStream_t s = c10::hip::getCurrentHIPStream((c10::DeviceIndex)device_).stream();
        int rc = ctlib_batch_run(b_, (void*)s);
        if (rc) throw std::runtime_error(std::string("ctlib_batch_run: ") + ctlib_batch_last_error());
    }
    void check() {
        c10::hip::HIPGuard guard((c10::DeviceIndex)device_);
        hipStream_t s = c10::hip::getCurrentHIPStream((c10::DeviceIndex)device_).stream();
        int rc = ctlib_batch_check(b_, (void*)s);
        if (rc) throw std::runtime_error(std::string("ctlib_batch_check: ") + ctlib_batch_last_error());
    }
    pybind11::dict info() const {
        int nitems = 0, bpc = 0, order = 0;
        ctlib_batch_info(b_, &nitems, &bpc, &order);
        pybind11::dict d;
        d["items"] = nitems;
        d["blocks_per_cu"] = bpc;
        d["order"] = order;
        d["slices"] = (int)geoms_.size();
        return d;
    }

  private:
    int device_;
    std::vector<ctlib_geom> geoms_;
    ctlib_batch* b_ = nullptr;
};

// Drops every cached plan; plans pinned by a graph capture stay alive (they are no longer
// looked up, a later call builds a fresh plan).
void clear_plan_cache() {
    std::lock_guard<std::mutex> lock(g_mu);
    g_cache.clear();
}

pybind11::dict plan_cache_info() {
    std::lock_guard<std::mutex> lock(g_mu);
    pybind11::dict d;
    d["entries"] = (int64_t)g_cache.size();
    d["pinned"] = (int64_t)g_pinned.size();
    d["created"] = g_created;
    d["capacity"] = (int64_t)kMaxPlans;
    return d;
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("projection", &projection, "CT projection (distance driven, HIP)");
    m.def("projection_t", &projection_t, "Transpose of CT projection (HIP)");
    m.def("backprojection_t", &backprojection_t, "Transpose of backprojection (HIP)");
    m.def("backprojection", &backprojection, "CT backprojection (HIP)");
    m.def("fbp", &fbp, "CT filtered backprojection with RL filter (HIP)");
    m.def("backprojection_t_true", &backprojection_t_true, "Exact adjoint of backprojection (HIP)");
    m.def("fbp_t", &fbp_t, "Exact adjoint of fbp (HIP)");
    m.def("backprojection_sv", &backprojection_sv, "Per-view backprojection, [B,V,H,W] (HIP)",
          pybind11::arg("projection"), pybind11::arg("options"), pybind11::arg("any_scan_type") = false);
    m.def("residual_step", &residual_step, "(A x, A^T (A x - m)) in one call (HIP)");
    m.def("residual_loss_step", &residual_loss_step,
          "(row sums of (A x - m)^2 [B, V] f64, A^T (A x - m)) in one call (HIP)");
    m.def("laplacian", &laplacian, "kNN graph weights as a sparse (N, N) tensor (HIP)");
    m.def("parse_options", &parse_options, "Parse an options vector");
    m.def("workspace_bytes", &workspace_bytes, "Scratch bytes of one op");
    m.def("clear_plan_cache", &clear_plan_cache, "Drop all cached geometry plans (graph-pinned plans stay alive)");
    m.def("plan_cache_info", &plan_cache_info, "Cached / pinned / created plan counts and the cache capacity");
    pybind11::class_<Batch>(m, "Batch")
        .def(pybind11::init<std::vector<torch::Tensor>, int64_t, bool>())
        .def("workspace_bytes", &Batch::workspace_bytes)
        .def("set_schedule", &Batch::set_schedule)
        .def("bind", &Batch::bind)
        .def("run", &Batch::run)
        .def("check", &Batch::check)
        .def("info", &Batch::info);
    m.def("set_plan_tuning", &set_plan_tuning, "Override the line->bin views per block of a cached plan");
    m.def("set_plan_chunks", &set_plan_chunks,
          "Override the line->bin line chunks and bin->pixel view chunks of a cached plan");
    m.def("set_plan_line_tables", &set_plan_line_tables,
          "Line->bin line tables of a cached plan: 0 prebuilt in the workspace, 1 built per block");
    m.def("timing_enable", &timing_enable, "Record HIP events around the library's kernel launches");
    m.def("timing_collect", &timing_collect, "Per-class kernel times since the last collect");
    m.attr("OP_PRJ") = (int)CTLIB_OP_PRJ;
    m.attr("OP_PRJ_T") = (int)CTLIB_OP_PRJ_T;
    m.attr("OP_BPRJ") = (int)CTLIB_OP_BPRJ;
    m.attr("OP_BPRJ_T") = (int)CTLIB_OP_BPRJ_T;
    m.attr("OP_BPRJ_T_TRUE") = (int)CTLIB_OP_BPRJ_T_TRUE;
    m.attr("OP_FBP") = (int)CTLIB_OP_FBP;
    m.attr("OP_FBP_T") = (int)CTLIB_OP_FBP_T;
    m.attr("OP_BPRJ_SV") = (int)CTLIB_OP_BPRJ_SV;
}
