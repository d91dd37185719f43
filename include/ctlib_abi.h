/*
 * C ABI of the MI355X (gfx950) distance-driven CT projector library
 * (libctlib_amd.so).  Plain C types only: no torch, no C++ exceptions.
 *
 * It replaces the reference's C++/CUDA dispatch layer:
 *   ctlib.cpp:15-99           projection / projection_t / backprojection /
 *                             backprojection_t / fbp  (scan_type switch)
 *   fan_ed_cuda.h:6-11, fan_ea_cuda.h:6-10, para_cuda.h:5-9
 *                             the per-geometry *_cuda host wrappers
 * The reference re-parses the device-resident options tensor on every call
 * (six .item() syncs, fan_ed_kernel.cu:1177-1181); here the options are parsed
 * once into an immutable plan that holds the per-view geometry tables on the
 * device, and ctlib_run only launches kernels on the caller's stream.
 *
 * Data layout (same as the reference's contiguous tensors, channel dim = 1):
 *   image    [batch][height][width]
 *   sinogram [batch][views][dets]     (bin 0 at the detector's negative end)
 */
#ifndef CTLIB_ABI_H
#define CTLIB_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ctlib_status {
    CTLIB_OK = 0,
    CTLIB_ERR_INVALID_ARG = 1,     /* bad pointer, size, or option value           */
    CTLIB_ERR_BAD_SCAN_TYPE = 2,   /* scan_type not in {0,1,2}; the reference exits */
    CTLIB_ERR_SHAPE = 3,           /* tensor shape does not match the options       */
    CTLIB_ERR_HIP = 4,             /* a HIP runtime call failed                      */
    CTLIB_ERR_UNSUPPORTED_DTYPE = 5,
    CTLIB_ERR_WORKSPACE = 6,       /* workspace missing or too small                 */
    CTLIB_ERR_GEOMETRY = 7         /* geometry the DD method cannot represent        */
} ctlib_status;

typedef enum ctlib_dtype { CTLIB_F32 = 0, CTLIB_F64 = 1 } ctlib_dtype;

typedef enum ctlib_op {
    CTLIB_OP_PRJ = 0,          /* projection        image -> sinogram  (ctlib.cpp:56-69) */
    CTLIB_OP_PRJ_T = 1,        /* projection_t      sinogram -> image  (ctlib.cpp:71-84) */
    CTLIB_OP_BPRJ = 2,         /* backprojection    sinogram -> image  (ctlib.cpp:30-43) */
    CTLIB_OP_BPRJ_T = 3,       /* backprojection_t  image -> sinogram  (ctlib.cpp:15-28),
                                  including fan-EA's extra dImg/c_k factor
                                  (fan_ea_kernel.cu:590-607)                            */
    CTLIB_OP_BPRJ_T_TRUE = 4,  /* exact adjoint of CTLIB_OP_BPRJ (differs from
                                  CTLIB_OP_BPRJ_T only for scan_type 1); no reference
                                  counterpart, used by autograd                           */
    CTLIB_OP_FBP = 5,          /* fbp               sinogram -> image  (ctlib.cpp:86-99) */
    CTLIB_OP_FBP_T = 6,        /* exact adjoint of CTLIB_OP_FBP; no reference counterpart */
    CTLIB_OP_BPRJ_SV = 7       /* backprojection_sv sinogram [batch][views][dets] -> per-view
                                  images [batch][views][height][width]: view v's term of
                                  CTLIB_OP_BPRJ (ctlib.cpp:45-54, fan_ed_kernel.cu:945-1161,
                                  1299-1331; the reference has it for scan_type 0 only,
                                  here every scan_type)                                   */
} ctlib_op;

/* Geometry as read from the options vector (README.md:56-76). */
typedef struct ctlib_geom {
    int32_t views, dets, width, height, scan_type;
    double dImg, dDet, Ang0, dAng, s2r, d2r, binshift;
} ctlib_geom;

typedef struct ctlib_plan ctlib_plan;

/* Parse an options vector of n entries (values widened to double from the
 * tensor dtype).  Index layout of fan_ed_kernel.cu:1165-1175 (fan, >= 11
 * entries) and para_kernel.cu:687-695 (parallel, >= 9 entries, binshift =
 * entry 8); scan_type is always the LAST entry (ctlib.cpp:18), so the README's
 * 9-entry parallel vector reads binshift = 2, as the reference does. */
int ctlib_parse_options(const double* opts, int n, ctlib_geom* out);

/* Build the plan (FP64 host tables, uploaded to `device`).  The plan is
 * immutable and may be shared by threads and streams on that device. */
int ctlib_plan_create(const ctlib_geom* geom, int device, ctlib_plan** out);
void ctlib_plan_destroy(ctlib_plan* plan);

/* Scratch bytes ctlib_run needs for (op, batch, dtype).  0 is possible. */
size_t ctlib_workspace_bytes(const ctlib_plan* plan, int op, int64_t batch, int dtype);

/* Run one operator on `stream` (a hipStream_t; NULL = legacy default stream).
 * Asynchronous: no host synchronisation, no allocation; `out` is fully
 * written (no caller memset).  `in`/`out` are device pointers laid out as
 * above; `ws` is device scratch of at least ctlib_workspace_bytes(). */
int ctlib_run(const ctlib_plan* plan, int op, int dtype, const void* in, void* out, int64_t batch,
              void* ws, size_t ws_bytes, void* stream);

/* Allocation-free variant of ctlib_run, with no counterpart in the reference: for
 * callers that cannot allocate device memory around a call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same ops, layouts,
 * stream semantics and results as ctlib_run, but the scratch comes from one fixed
 * device array of CTLIB_NOALLOC_SCRATCH_BYTES per GPU inside the library (held by
 * every GPU the library runs on until the process ends), and the call makes no
 * hipMalloc / hipFree and no other allocation.  When the automatic launch split of
 * (op, batch) needs more scratch than that, fewer chunks are used; when even an
 * unsplit launch does not fit, it returns CTLIB_ERR_WORKSPACE without launching.
 * Every call on one GPU uses that GPU's array and nothing in it carries over between
 * calls, so calls that overlap in time on one GPU (another thread or another stream)
 * race: serialise them.  The array's address on a device is resolved when the first
 * plan is created there. */
#define CTLIB_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ctlib_run_noalloc(const ctlib_plan* plan, int op, int dtype, const void* in, void* out, int64_t batch,
                      void* stream);

/* One least-squares data-term step, the unit of work of an unrolled
 * reconstruction:  y = A x  and  grad = A^T (A x - m).  `y` may be NULL.
 * No reference counterpart (the reference has no fused path). */
int ctlib_run_residual(const ctlib_plan* plan, int dtype, const void* x, const void* m, void* y,
                       void* grad, int64_t batch, void* ws, size_t ws_bytes, void* stream);
size_t ctlib_residual_workspace_bytes(const ctlib_plan* plan, int64_t batch, int dtype);
/* As ctlib_run_residual, and row_sq[b * views + v] = sum over the bins of row (b, v) of
 * (y - m)^2 in FP64, y rounded to the data type first, so that 0.5 * sum(row_sq) is the
 * least-squares loss without another pass over y and m.  row_sq: device array of
 * batch * views doubles, or NULL.  Same workspace as ctlib_run_residual.  No reference
 * counterpart. */
int ctlib_run_residual_loss(const ctlib_plan* plan, int dtype, const void* x, const void* m, void* y,
                            void* grad, double* row_sq, int64_t batch, void* ws, size_t ws_bytes,
                            void* stream);

/* Launch-shape override chosen by the autotuner: views per block of the
 * line->bin kernel (4 or 8; 0 = automatic; 8 applies when D < 768).  Set it before the plan is
 * used from several threads; it applies to later ctlib_run calls. */
int ctlib_plan_set_tuning(ctlib_plan* plan, int fwd_views_per_block);

/* Launch-shape overrides of work splitting, for batches too small to fill the
 * GPU: the line->bin kernel (projection, backprojection_t, fbp_t) splits each
 * sweep axis's image lines into `line_chunks` chunks, the bin->pixel kernel
 * (projection_t, backprojection, fbp) its views into `view_chunks` chunks; the
 * partial sinograms / images are summed afterwards.  0 = automatic (about 4096
 * blocks per launch), 1 = no split, at most 64 (capped at the lines / views of
 * each axis).  The workspace size depends on them: set them before
 * ctlib_workspace_bytes and before the plan is used from several threads.
 * ctlib_run_residual splits lines too (its forward is then not fused with the
 * residual tables: y is summed from the partials and then y - m is tabulated).
 * No reference counterpart (launch shape only). */
int ctlib_plan_set_chunks(ctlib_plan* plan, int line_chunks, int view_chunks);

/* Where the unweighted line->bin kernel (projection, and the forward of
 * ctlib_run_residual / ctlib_run_residual_loss) gets the cumulative table of
 * each image line.  0 (default): one pass builds every line's table into the
 * workspace first (batch * (H + W) * (256*ceil(max(H, W)/256) + 2) * 16 bytes
 * at most), and the kernel copies them into LDS; 1: each block builds the
 * tables of the lines it sweeps (no table workspace, more arithmetic per
 * block).  Same results up to FP64 summation order.  ctlib_run_noalloc uses 1
 * by itself when the tables do not fit its scratch.  The workspace size depends
 * on it: set it before ctlib_workspace_bytes / ctlib_residual_workspace_bytes
 * and before the plan is used from several threads.  No reference counterpart
 * (launch shape only). */
int ctlib_plan_set_line_tables(ctlib_plan* plan, int per_block);

/* Geometry queries for host code that mirrors the plan (tests, autograd). */
int ctlib_plan_geom(const ctlib_plan* plan, ctlib_geom* out);
int ctlib_plan_view_info(const ctlib_plan* plan, int view, int* xmode, int* rev);

/* Host copy of one plan table, built without touching a GPU (for host-side
 * tooling and tests).  table: 0 view params (64-byte records), 1 bin edges
 * [V][D+1] f64, 2 projection factors [V][D] f64, 3 backprojection_t factors
 * [V][D] f64, 4 line coefficients [V][max(H,W)] (64-byte records), 5 X-axis
 * view list int32, 6 Y-axis view list int32, 7 RL taps [2D-1] f64, 8 fbp
 * pre-weight [D] f64, 9 fan-EA arctangent polynomial as 8 f32: c0..c5 of
 * atan(t) ~ t (c0 + c1 t^2 + ... + c5 t^10) on |t| <= T, then T, then 1 if the
 * kernels use it (else they call atanf), 10 the view lookup composed with each
 * line's pixel-edge map [V][max(H,W)][4] f32 {64 lk0 s, lk0 a + lk1, 64 lk2 s,
 * lk2 a + lk3}, 11 each line's pixel weight (1/s) / ((w0 + w1 p)^2 + w2) as
 * [V][max(H,W)][4] f32 {64 c w1, c w1, c w0, c^2 w2}, c = sqrt(s): the weight at
 * pixel p = j + 64 k is 1 / (u^2 + c^2 w2) with u = k (64 c w1) + j c w1 + c w0.
 * With dst == NULL only *bytes is set. */
int ctlib_plan_export(const ctlib_geom* geom, int table, void* dst, size_t* bytes);

/* Heterogeneous batch: slices of different geometry (image size, views, bins,
 * scan_type), each with one image x_s and one measured sinogram m_s, run as
 * ONE persistent fused launch computing y_s = A_s x_s and
 * grad_s = A_s^T (A_s x_s - m_s) for every slice.  Work items (line copies,
 * line tables, projection groups, projection_t tiles, axis combines) are dequeued
 * from one atomic counter; per-slice counters order the phases.  ctlib_batch_run
 * allocates nothing, synchronises nothing and records no memset, so it may be
 * captured in a HIP graph.  Slices: width, height <= 1024, dets <= 1536 (a batch
 * with any slice beyond 512 x 512 or 1279 bins runs a kernel variant whose
 * line tables hold up to 1024 pixels, for every slice of that batch).
 * No reference counterpart (the reference runs one geometry per call). */
typedef struct ctlib_batch ctlib_batch;
int ctlib_batch_create(const ctlib_geom* geoms, int nslices, int device, int dtype, ctlib_batch** out);
void ctlib_batch_destroy(ctlib_batch* batch);
size_t ctlib_batch_workspace_bytes(const ctlib_batch* batch);
/* Schedule knobs chosen by the autotuner: resident blocks per CU (1..8) and
 * item order within a phase (0 longest first, 1 slice order). */
int ctlib_batch_set_schedule(ctlib_batch* batch, int blocks_per_cu, int order);
/* Bind per-slice device pointers (x, m, grad required; y optional) and the
 * workspace; uploads the slice records (synchronous).  Rebind when any changes. */
int ctlib_batch_bind(ctlib_batch* batch, const void* const* x, const void* const* m, void* const* y,
                     void* const* grad, void* ws, size_t ws_bytes);
int ctlib_batch_run(const ctlib_batch* batch, void* stream);
int ctlib_batch_info(const ctlib_batch* batch, int* nitems, int* blocks_per_cu, int* order);
/* Waits for `stream` and reports whether the last run flagged a dependency-wait timeout. */
int ctlib_batch_check(const ctlib_batch* batch, void* stream);
const char* ctlib_batch_last_error(void);

/* kNN graph weights of the MAGIC network (laplacian_cuda_kernel.cu:46-128,
 * ctlib.cpp:101-104): for n points x [n][d] (dtype f32 or f64, device memory),
 * the k nearest points of each point by squared distance (itself included),
 * cols [n][k] int64 in ascending distance (ties by column), and their weights
 * exp(-dist / median) normalised per point, vals [n][k] float32, where median
 * is the lower median of all n*k selected distances.  Row i of the sparse
 * matrix holds cols[i] / vals[i].  Supported: any d, k <= 4096 (k <= n).
 * Asynchronous on `stream`. */
size_t ctlib_laplacian_workspace_bytes(int64_t n, int d, int k, int dtype);
int ctlib_laplacian(const void* x, int64_t n, int d, int k, int dtype, int64_t* cols, float* vals, void* ws,
                    size_t ws_bytes, void* stream);

/* Kernel timing.  While enabled, ctlib_run / ctlib_run_residual record a HIP
 * event pair on the launch stream around each launch, grouped in classes:
 * 0 line->bin DD kernel, 1 bin->pixel DD kernel, 2 ramp GEMM, 3 auxiliary
 * (line transposes, bin tables, axis combine), 4 per-view backprojection
 * (CTLIB_OP_BPRJ_SV; its algo_bytes are es*B*V*(H*W + D), the images written
 * plus the sinogram read); classes 5-7 are unused.  ctlib_timing_collect waits for
 * the recorded events, sums them per class and clears the record.
 * algo_bytes: sweep-model bytes of the DD kernels, es*B*V*(H*W + D) per launch;
 * flops: 2*B*V*D*D per ramp launch.  No reference counterpart. */
typedef struct ctlib_kernel_stats {
    int64_t launches[8];
    double ms[8];
    double algo_bytes[8];
    double flops[8];
} ctlib_kernel_stats;
int ctlib_timing_enable(int on);
int ctlib_timing_collect(ctlib_kernel_stats* out);

/* Last error text set on the calling thread (never NULL). */
const char* ctlib_last_error(void);
const char* ctlib_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif /* CTLIB_ABI_H */
