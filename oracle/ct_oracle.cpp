// FP64 CPU oracle for CTLIB's distance-driven (DD) projector family.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (ctlib_amd/) may link, load
// or call this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only as the checker / CPU baseline.
//
// Two independent implementations live here:
//   * ctor_apply  - a double-precision restatement of the reference's merge
//     loops (per view, per 256-bin detector chunk, per image line), keeping the
//     reference's branch order and its strict/non-strict comparisons.  It
//     computes the intended values of the reference kernels (no shared-memory
//     races, no fp32 angle product, closed-form pixel edges).
//       projection        fan_ed_kernel.cu:23-254   fan_ea_kernel.cu:22-240   para_kernel.cu:7-170
//       projection_t      fan_ed_kernel.cu:256-499  fan_ea_kernel.cu:242-472  para_kernel.cu:172-336
//       backprojection_t  fan_ed_kernel.cu:501-705  fan_ea_kernel.cu:474-698  para_kernel.cu:338-499
//       backprojection    fan_ed_kernel.cu:707-923  fan_ea_kernel.cu:700-901  para_kernel.cu:501-663
//       fbp               fan_ed_kernel.cu:1333-1382 fan_ea_kernel.cu:1059-1107 para_kernel.cu:809-849
//       backprojection_sv fan_ed_kernel.cu:945-1161, 1299-1331 (per-view images; the
//                         same merge for fan-EA and parallel beam as an extension)
//       rlfilter          fan_ed_kernel.cu:925-943  fan_ea_kernel.cu:903-921  para_kernel.cu:665-683
//       options/dispatch  ctlib.cpp:15-99, fan_ed_kernel.cu:1165-1175, para_kernel.cu:687-695
//   * ctor_brute  - an O(views*lines*pixels*bins) interval-intersection
//     implementation with its own ray/axis intersection arithmetic and no
//     merge, chunking or start-pixel search.  It exists to pin ctor_apply.
//
// Conventions (shared with the reference):
//   image  [B][H][W] row-major, row 0 at the top (y = (H/2 - r - 1/2) dImg)
//   sino   [B][V][D] row-major, bin j spans [(j - D/2), (j + 1 - D/2)] dDet-units
//   op codes: 0 projection, 1 projection_t, 2 backprojection,
//             3 backprojection_t (reference form), 4 backprojection_t true adjoint
//             (differs from 3 only for scan_type 1), 5 fbp,
//             6 backprojection_sv (sinogram [B][V][D] -> images [B][V][H][W]).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { OP_PRJ = 0, OP_PRJ_T = 1, OP_BPRJ = 2, OP_BPRJ_T = 3, OP_BPRJ_T_TRUE = 4, OP_FBP = 5, OP_BPRJ_SV = 6 };
enum { ST_ED = 0, ST_EA = 1, ST_PARA = 2 };
enum { E_OK = 0, E_ARG = 1, E_SCAN = 2 };

struct Geom {
    int V, D, W, H, scan;
    double dImg, dDet, Ang0, dAng, s2r, d2r, bs;
};

// Option indexing of ctlib.cpp:18 (scan_type = last entry) and
// fan_ed_kernel.cu:1165-1175 / para_kernel.cu:687-695 (geometry entries).
int parse(const double* o, int n, Geom* g) {
    if (!o || n < 1) return E_ARG;
    double st = o[n - 1];
    if (!(st == 0.0 || st == 1.0 || st == 2.0)) return E_SCAN;
    g->scan = (int)st;
    int need = g->scan == ST_PARA ? 9 : 11;
    if (n < need) return E_ARG;
    for (int i = 0; i < 4; i++)
        if (!(o[i] >= 1.0) || o[i] != std::floor(o[i]) || o[i] > 1e6) return E_ARG;
    g->V = (int)o[0]; g->D = (int)o[1]; g->W = (int)o[2]; g->H = (int)o[3];
    g->dImg = o[4]; g->dDet = o[5]; g->Ang0 = o[6]; g->dAng = o[7];
    if (g->scan == ST_PARA) { g->s2r = g->d2r = 0.0; g->bs = o[8]; }
    else { g->s2r = o[8]; g->d2r = o[9]; g->bs = o[10]; }
    if (!(g->dImg > 0) || !(g->dDet > 0)) return E_ARG;
    if (g->scan != ST_PARA && !(g->s2r > 0)) return E_ARG;
    return E_OK;
}

// fan_ed_kernel.cu:8-10 / :13-15
inline double map_x(double sx, double sy, double px, double py) { return (sx * py - sy * px) / (py - sy); }
inline double map_y(double sx, double sy, double px, double py) { return (sy * px - sx * py) / (px - sx); }

struct View {
    double ang, cs, sn, sx, sy;
    bool xmode, rev;
    std::vector<double> e;   // D+1 axis edges, internal (merge) order
    std::vector<double> cb;  // D bin cosines c_k (fan)
};

// Per-view setup: fan_ed_kernel.cu:58-96 (and twins fan_ea_kernel.cu:55-85, para_kernel.cu:36-60),
// with the angle product in double.
void view_setup(const Geom& g, int v, View& w) {
    const double PI = std::acos(-1.0);
    double ang = (double)v * g.dAng + g.Ang0;
    w.ang = ang;
    double err = std::fabs(ang - std::round(ang / PI) * PI) * 4 / PI;
    w.xmode = err <= 1;
    double err2 = (ang - std::floor(ang / 2 / PI) * 2 * PI) * 4 / PI;
    w.rev = err2 >= 3 && err2 < 7;
    w.cs = std::cos(ang);
    w.sn = std::sin(ang);
    w.sx = -w.sn * g.s2r;
    w.sy = w.cs * g.s2r;
    const int D = g.D;
    const double hD = (double)D / 2;
    w.e.assign(D + 1, 0.0);
    for (int i = 0; i <= D; i++) {
        double e;
        if (g.scan == ST_ED) {
            double vd = g.s2r / (g.s2r + g.d2r) * g.dDet;
            double vs = g.s2r / (g.s2r + g.d2r) * g.bs;
            double u = w.rev ? (hD - i) * vd + vs : (i - hD) * vd + vs;
            double px = u * w.cs, py = u * w.sn;
            e = w.xmode ? map_x(w.sx, w.sy, px, py) : map_y(w.sx, w.sy, px, py);
        } else if (g.scan == ST_EA) {
            double beta = w.rev ? (hD - i) * g.dDet + g.bs + ang : (i - hD) * g.dDet + g.bs + ang;
            e = w.xmode ? w.sx + std::sin(beta) / std::cos(beta) * w.sy
                        : w.sy + std::cos(beta) / std::sin(beta) * w.sx;
        } else {
            double t = w.rev ? (hD - i) * g.dDet + g.bs : (i - hD) * g.dDet + g.bs;
            e = w.xmode ? t / w.cs : t / w.sn;
        }
        w.e[i] = e;
    }
    w.cb.assign(D, 0.0);
    if (g.scan != ST_PARA) {
        // fan_ed_kernel.cu:146-154 (X) and :235-243 (Y)
        auto kap = [&](double e) {
            return w.xmode ? std::fabs(w.sy) / std::sqrt((e - w.sx) * (e - w.sx) + w.sy * w.sy)
                           : std::fabs(w.sx) / std::sqrt((e - w.sy) * (e - w.sy) + w.sx * w.sx);
        };
        for (int k = 0; k < D; k++) w.cb[k] = (kap(w.e[k]) + kap(w.e[k + 1])) / 2;
    }
}

// Line geometry: a line is an image row (X-mode) or column (Y-mode).
struct LineGeo {
    int npix;       // pixels along the line
    double lc;      // line centre coordinate (y of row / x of column)
    double p0;      // first pixel-edge coordinate along the line
};

inline LineGeo line_geo(const Geom& g, const View& w, int l) {
    LineGeo L;
    if (w.xmode) { L.npix = g.W; L.lc = (g.H / 2.0 - l - 0.5) * g.dImg; L.p0 = -g.W / 2.0 * g.dImg; }
    else         { L.npix = g.H; L.lc = (l - g.W / 2.0 + 0.5) * g.dImg; L.p0 = -g.H / 2.0 * g.dImg; }
    return L;
}

// Axis coordinate of pixel edge p of a line (closed form of the reference's i0x += dImage).
inline double pix_edge(const Geom& g, const View& w, const LineGeo& L, int p) {
    double t = L.p0 + p * g.dImg;
    if (g.scan == ST_PARA) return w.xmode ? w.sn / w.cs * L.lc + t : t + w.cs / w.sn * L.lc;
    return w.xmode ? map_x(w.sx, w.sy, t, L.lc) : map_y(w.sx, w.sy, L.lc, t);
}

// Pixel-centre weight of backprojection: fan_ed_kernel.cu:17-21 (ED), fan_ea_kernel.cu:17-20 (EA, applied as 1/L^2).
inline double pix_weight(const Geom& g, const View& w, const LineGeo& L, int p) {
    if (g.scan == ST_PARA) return 1.0;
    double t = L.p0 + (p + 0.5) * g.dImg;
    double x = w.xmode ? t : L.lc, y = w.xmode ? L.lc : t;
    if (g.scan == ST_ED) {
        double r = g.s2r, d = (w.sx * x + w.sy * y) / r;
        return r * r / ((r - d) * (r - d));
    }
    double l2 = (w.sx - x) * (w.sx - x) + (w.sy - y) * (w.sy - y);
    return 1.0 / l2;
}

// First pixel of a line reached by the chunk-start ray (fan_ed_kernel.cu:108-110,197-199;
// fan_ea_kernel.cu:98-100; para_kernel.cu:66-67,126-127), clamped at 0.
int start_pixel(const Geom& g, const View& w, const LineGeo& L, int c0) {
    double dP0 = w.e[c0];
    double f;
    if (g.scan == ST_ED) {
        const double hD = (double)g.D / 2;
        double vd = g.s2r / (g.s2r + g.d2r) * g.dDet;
        double vs = g.s2r / (g.s2r + g.d2r) * g.bs;
        double u = w.rev ? (hD - c0) * vd + vs : (c0 - hD) * vd + vs;
        double d0x = u * w.cs, d0y = u * w.sn;
        f = w.xmode ? ((L.lc - w.sy) / (d0y - w.sy) * (d0x - w.sx) + w.sx - L.p0) / g.dImg
                    : ((L.lc - w.sx) / (d0x - w.sx) * (d0y - w.sy) + w.sy - L.p0) / g.dImg;
    } else if (g.scan == ST_EA) {
        f = w.xmode ? ((L.lc - w.sy) / (-w.sy) * (dP0 - w.sx) + w.sx - L.p0) / g.dImg
                    : ((L.lc - w.sx) / (-w.sx) * (dP0 - w.sy) + w.sy - L.p0) / g.dImg;
    } else {
        double pb = w.xmode ? w.sn / w.cs * L.lc + L.p0 : L.p0 + w.cs / w.sn * L.lc;
        f = (dP0 - pb) / g.dImg;
    }
    f = std::floor(f);
    if (!(f > 0)) return 0;           // max(idx0, 0); NaN -> 0
    if (f > L.npix) return L.npix;    // loop would not run
    return (int)f;
}

inline double& pix_ref(const Geom& g, const View& w, double* img, int l, int p) {
    return w.xmode ? img[(size_t)l * g.W + p] : img[(size_t)(g.H - 1 - p) * g.W + l];
}
inline double pix_val(const Geom& g, const View& w, const double* img, int l, int p) {
    return w.xmode ? img[(size_t)l * g.W + p] : img[(size_t)(g.H - 1 - p) * g.W + l];
}

const int CHUNK_DEFAULT = 256;  // BLOCK_DIM, fan_ed_kernel.cu:4

// Sinogram-output merge (projection and backprojection_t) for one (view, chunk, line).
// `wmode` 0: projection (ov * x / w_k, or * dImg/dDet for para)
//         1: backprojection_t (ov * x / s_p * W(p))
void merge_to_bins(const Geom& g, const View& w, int c0, int nb, int l, const double* img,
                   double* prj, int wmode) {
    LineGeo L = line_geo(g, w, l);
    const double* dP = &w.e[c0 + 1];   // dPoint[t] = right edge of bin c0+t
    double dP0 = w.e[c0];
    int i0 = start_pixel(g, w, L, c0);
    if (i0 >= L.npix) return;
    double acc = 0;
    double prepix = pix_edge(g, w, L, i0);
    double pre = std::max(prepix, dP0);
    double pixb = pix_edge(g, w, L, i0 + 1);
    double detb = dP[0];
    int d = 0, i = i0;
    const double dinterval = g.dImg / g.dDet;
    auto term = [&](double ov) {
        double xv = pix_val(g, w, img, l, i);
        if (wmode == 0) {
            if (g.scan == ST_PARA) return ov * xv * dinterval;
            double coef = d == 0 ? dP[0] - dP0 : dP[d] - dP[d - 1];
            return ov * xv / coef;
        }
        if (g.scan == ST_PARA) return ov * xv / g.dImg;
        return ov * xv / (pixb - prepix) * pix_weight(g, w, L, i);
    };
    while (i < L.npix && d < nb) {
        if (detb <= pre) {
            d++;
            if (d < nb) detb = dP[d];
        } else if (pixb <= pre) {
            i++;
            prepix = pixb;
            pixb = pix_edge(g, w, L, i + 1);
        } else if (pixb < detb) {
            acc += term(pixb - pre);
            pre = pixb;
            i++;
            prepix = pixb;
            pixb = pix_edge(g, w, L, i + 1);
        } else {
            acc += term(detb - pre);
            pre = detb;
            prj[d] += acc;
            acc = 0;
            d++;
            if (d < nb) detb = dP[d];
        }
    }
    if (acc != 0) prj[d] += acc;
}

// Image-output merge (projection_t and backprojection) for one (view, chunk, line).
// `gb` is the per-bin value already in internal order for this chunk:
//   projection_t:   y * dImg / (c_k * w_k)   (fan)  or  y * dImg/dDet (para, applied per term)
//   backprojection: y, divided by s_p per term and weighted by W(p) per finished pixel
void merge_to_pixels(const Geom& g, const View& w, int c0, int nb, int l, const double* gb,
                     double* img, int wmode) {
    LineGeo L = line_geo(g, w, l);
    const double* dP = &w.e[c0 + 1];
    double dP0 = w.e[c0];
    int i0 = start_pixel(g, w, L, c0);
    if (i0 >= L.npix) return;
    double acc = 0;
    double prepix = pix_edge(g, w, L, i0);
    double pre = std::max(prepix, dP0);
    double pixb = pix_edge(g, w, L, i0 + 1);
    double detb = dP[0];
    int d = 0, i = i0;
    const double dinterval = g.dImg / g.dDet;
    auto term = [&](double ov) {
        if (wmode == 0) return g.scan == ST_PARA ? ov * gb[d] * dinterval : ov * gb[d];
        return g.scan == ST_PARA ? ov * gb[d] / g.dImg : ov * gb[d] / (pixb - prepix);
    };
    auto flush = [&]() {
        double v = acc;
        if (wmode == 1 && g.scan == ST_ED) v *= pix_weight(g, w, L, i);
        if (wmode == 1 && g.scan == ST_EA) v *= pix_weight(g, w, L, i);  // /= |s-p|^2
        pix_ref(g, w, img, l, i) += v;
        acc = 0;
    };
    while (i < L.npix && d < nb) {
        if (detb <= pre) {
            d++;
            if (d < nb) detb = dP[d];
        } else if (pixb <= pre) {
            i++;
            prepix = pixb;
            pixb = pix_edge(g, w, L, i + 1);
        } else if (pixb <= detb) {
            acc += term(pixb - pre);
            pre = pixb;
            flush();
            i++;
            prepix = pixb;
            pixb = pix_edge(g, w, L, i + 1);
        } else {
            acc += term(detb - pre);
            pre = detb;
            d++;
            if (d < nb) detb = dP[d];
        }
    }
    if (acc != 0) flush();
}

inline int out_index(const Geom& g, const View& w, int k) { return w.rev ? g.D - 1 - k : k; }

void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

// Sinogram-output ops: projection (wmode 0) and backprojection_t (wmode 1, 2 = EA-quirk off).
void run_to_sino(const Geom& g, const std::vector<View>& views, const double* in, double* out, int B,
                 int chunk, int op) {
    const int V = g.V, D = g.D;
    #pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int b = 0; b < B; b++) {
        for (int v = 0; v < V; v++) {
            const View& w = views[v];
            const double* img = in + (size_t)b * g.H * g.W;
            double* y = out + ((size_t)b * V + v) * D;
            std::vector<double> prj(D, 0.0);
            int nlines = w.xmode ? g.H : g.W;
            for (int c0 = 0; c0 < D; c0 += chunk) {
                int nb = std::min(chunk, D - c0);
                for (int l = 0; l < nlines; l++)
                    merge_to_bins(g, w, c0, nb, l, img, &prj[c0], op == OP_PRJ ? 0 : 1);
            }
            for (int k = 0; k < D; k++) {
                double val = prj[k];
                if (op == OP_PRJ && g.scan != ST_PARA) val = val * g.dImg / w.cb[k];
                if (op == OP_BPRJ_T && g.scan == ST_EA) val = val * g.dImg / w.cb[k];  // fan_ea_kernel.cu:590-607
                y[out_index(g, w, k)] = val;
            }
        }
    }
}

// Image-output ops: projection_t (wmode 0) and backprojection (wmode 1).  Lines are
// independent within a mode, so each line is owned by one thread and receives its
// views in ascending order: deterministic for any thread count.
void run_to_image(const Geom& g, const std::vector<View>& views, const double* in, double* out, int B,
                  int chunk, int op) {
    const int V = g.V, D = g.D;
    std::vector<int> vx, vy;
    for (int v = 0; v < V; v++) (views[v].xmode ? vx : vy).push_back(v);
    std::vector<double> gb((size_t)V * D);
    for (int b = 0; b < B; b++) {
        const double* y = in + (size_t)b * V * D;
        double* img = out + (size_t)b * g.H * g.W;
        std::fill(img, img + (size_t)g.H * g.W, 0.0);
        for (int v = 0; v < V; v++) {
            const View& w = views[v];
            for (int k = 0; k < D; k++) {
                double val = y[(size_t)v * D + out_index(g, w, k)];
                if (op == OP_PRJ_T && g.scan != ST_PARA) {
                    // fan_ed_kernel.cu:343-357: y * dImg / c_k / w_k
                    val = val * g.dImg / w.cb[k];
                    val = val / (w.e[k + 1] - w.e[k]);
                }
                gb[(size_t)v * D + k] = val;
            }
        }
        for (int pass = 0; pass < 2; pass++) {
            const std::vector<int>& vl = pass == 0 ? vx : vy;
            int nlines = pass == 0 ? g.H : g.W;
            #pragma omp parallel for schedule(dynamic, 4)
            for (int l = 0; l < nlines; l++) {
                for (int v : vl) {
                    const View& w = views[v];
                    for (int c0 = 0; c0 < D; c0 += chunk) {
                        int nb = std::min(chunk, D - c0);
                        merge_to_pixels(g, w, c0, nb, l, &gb[(size_t)v * D + c0], img,
                                        op == OP_PRJ_T ? 0 : 1);
                    }
                }
            }
        }
    }
}

// Per-view backprojection (backprojection_sv, fan_ed_kernel.cu:945-1161): the merge
// of backprojection, written to image (b, v) instead of summed over views
// (fan_ed_kernel.cu:1042,1128 index image[channel][view]).
void run_sv(const Geom& g, const std::vector<View>& views, const double* in, double* out, int B, int chunk) {
    const int V = g.V, D = g.D;
    const size_t npx = (size_t)g.H * g.W;
    #pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int b = 0; b < B; b++) {
        for (int v = 0; v < V; v++) {
            const View& w = views[v];
            double* img = out + ((size_t)b * V + v) * npx;
            std::fill(img, img + npx, 0.0);
            std::vector<double> gb(D);
            for (int k = 0; k < D; k++) gb[k] = in[((size_t)b * V + v) * D + out_index(g, w, k)];
            const int nlines = w.xmode ? g.H : g.W;
            for (int l = 0; l < nlines; l++)
                for (int c0 = 0; c0 < D; c0 += chunk)
                    merge_to_pixels(g, w, c0, std::min(chunk, D - c0), l, &gb[c0], img, 1);
        }
    }
}

// RL taps h(n), n = -(D-1)..D-1 (fan_ed_kernel.cu:925-943).
void rl_taps(int D, double delta, double* h) {
    const double PI = std::acos(-1.0);
    for (int t = 0; t < 2 * D - 1; t++) {
        int x = t - D + 1;
        if (std::abs(x) % 2 == 1) h[t] = -1.0 / (PI * PI * x * x * delta * delta);
        else if (x == 0) h[t] = 1.0 / (4 * delta * delta);
        else h[t] = 0.0;
    }
}

// Pre-weight and ramp spacing of fbp (fan_ed_kernel.cu:1349-1355,1365; fan_ea_kernel.cu:1075-1080,1090;
// para_kernel.cu:823,833).
void fbp_preweight(const Geom& g, double* pw, double* delta) {
    const int D = g.D;
    for (int j = 0; j < D; j++) {
        double r = -D / 2.0 + 0.5 + j;
        if (g.scan == ST_ED) {
            double vd = g.dDet * g.s2r / (g.s2r + g.d2r);
            double t = r * vd;
            pw[j] = g.s2r / std::sqrt(g.s2r * g.s2r + t * t) * vd;
        } else if (g.scan == ST_EA) {
            pw[j] = std::cos(r * g.dDet) * g.s2r * g.dDet;
        } else {
            pw[j] = g.dDet;
        }
    }
    *delta = g.scan == ST_ED ? g.dDet * g.s2r / (g.s2r + g.d2r) : g.dDet;
}

// conv2d(q, h, padding = (0, D-1)) as cross-correlation (fan_ed_kernel.cu:1368).
void ramp_rows(const Geom& g, const double* y, double* q_out, int rows) {
    const int D = g.D;
    std::vector<double> pw(D), h(2 * D - 1);
    double delta;
    fbp_preweight(g, pw.data(), &delta);
    rl_taps(D, delta, h.data());
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
        const double* yr = y + (size_t)r * D;
        double* qr = q_out + (size_t)r * D;
        for (int j = 0; j < D; j++) {
            double s = 0;
            for (int m = 0; m < D; m++) s += yr[m] * pw[m] * h[m - j + D - 1];
            qr[j] = s;
        }
    }
}

std::vector<View> all_views(const Geom& g) {
    std::vector<View> views(g.V);
    for (int v = 0; v < g.V; v++) view_setup(g, v, views[v]);
    return views;
}

// ---------------------------------------------------------------------------------
// Brute force: independent ray/axis intersections and all-pairs interval overlap.
// ---------------------------------------------------------------------------------

// Intersection with the sweep axis (y=0 in X-mode, x=0 in Y-mode) of the line through
// point (px,py) with direction (dx,dy).
inline double axis_hit(bool xmode, double px, double py, double dx, double dy) {
    if (xmode) { double lam = -py / dy; return px + lam * dx; }
    double lam = -px / dx;
    return py + lam * dy;
}

struct BView {
    bool xmode, rev;
    double sx, sy, cs, sn, ang;
    std::vector<double> e;  // D+1 edges indexed by the OUTPUT bin boundary j (bin j = [e_j, e_{j+1}] or reversed)
};

void bview(const Geom& g, int v, BView& w) {
    const double PI = std::acos(-1.0);
    double ang = (double)v * g.dAng + g.Ang0;
    w.ang = ang;
    // Axis choice and orientation are definitions of the reference (fan_ed_kernel.cu:63,75-76).
    w.xmode = std::fabs(ang - std::round(ang / PI) * PI) * 4 / PI <= 1;
    double r = (ang - std::floor(ang / 2 / PI) * 2 * PI) * 4 / PI;
    w.rev = r >= 3 && r < 7;
    w.cs = std::cos(ang); w.sn = std::sin(ang);
    w.sx = -w.sn * g.s2r; w.sy = w.cs * g.s2r;
    w.e.assign(g.D + 1, 0.0);
    for (int j = 0; j <= g.D; j++) {
        double q = j - g.D / 2.0;  // boundary j in detector units
        if (g.scan == ST_ED) {
            double mag = g.s2r / (g.s2r + g.d2r);
            double u = (q * g.dDet + g.bs) * mag;                 // point on the virtual detector
            double px = u * w.cs, py = u * w.sn;
            w.e[j] = axis_hit(w.xmode, w.sx, w.sy, px - w.sx, py - w.sy);
        } else if (g.scan == ST_EA) {
            double beta = ang + q * g.dDet + g.bs;                // ray direction (sin b, -cos b)
            w.e[j] = axis_hit(w.xmode, w.sx, w.sy, std::sin(beta), -std::cos(beta));
        } else {
            double t = q * g.dDet + g.bs;
            w.e[j] = axis_hit(w.xmode, t * w.cs, t * w.sn, -w.sn, w.cs);
        }
    }
}

// Axis coordinate of an image point projected along its ray.
inline double bproj_point(const Geom& g, const BView& w, double x, double y) {
    if (g.scan == ST_PARA) return axis_hit(w.xmode, x, y, -w.sn, w.cs);
    return axis_hit(w.xmode, w.sx, w.sy, x - w.sx, y - w.sy);
}

void brute(const Geom& g, int op, const double* in, double* out, int B) {
    const int V = g.V, D = g.D, H = g.H, W = g.W;
    std::vector<BView> views(V);
    for (int v = 0; v < V; v++) bview(g, v, views[v]);
    bool to_sino = op == OP_PRJ || op == OP_BPRJ_T || op == OP_BPRJ_T_TRUE;
    if (to_sino) std::fill(out, out + (size_t)B * V * D, 0.0);
    else if (op == OP_BPRJ_SV) std::fill(out, out + (size_t)B * V * H * W, 0.0);
    else std::fill(out, out + (size_t)B * H * W, 0.0);
    for (int v = 0; v < V; v++) {
        const BView& w = views[v];
        // bin j interval on the axis, its cosine factor and width
        std::vector<double> lo(D), hi(D), cb(D);
        for (int j = 0; j < D; j++) {
            lo[j] = std::min(w.e[j], w.e[j + 1]);
            hi[j] = std::max(w.e[j], w.e[j + 1]);
            if (g.scan != ST_PARA) {
                auto cosr = [&](double e) {  // |cos| of the angle between the ray to axis point e and the axis normal
                    double dx = w.xmode ? e - w.sx : -w.sx, dy = w.xmode ? -w.sy : e - w.sy;
                    double nrm = std::sqrt(dx * dx + dy * dy);
                    return w.xmode ? std::fabs(dy) / nrm : std::fabs(dx) / nrm;
                };
                cb[j] = 0.5 * (cosr(w.e[j]) + cosr(w.e[j + 1]));
            }
        }
        int nlines = w.xmode ? H : W, npix = w.xmode ? W : H;
        for (int l = 0; l < nlines; l++) {
            for (int p = 0; p < npix; p++) {
                int r, c;
                double x0, x1, y0, y1;  // pixel edge points along the line at the line centre
                if (w.xmode) {
                    r = l; c = p;
                    double yc = (H / 2.0 - r - 0.5) * g.dImg;
                    x0 = (c - W / 2.0) * g.dImg; x1 = x0 + g.dImg; y0 = y1 = yc;
                } else {
                    c = l; r = H - 1 - p;
                    double xc = (c - W / 2.0 + 0.5) * g.dImg;
                    y0 = (p - H / 2.0) * g.dImg; y1 = y0 + g.dImg; x0 = x1 = xc;
                }
                double a0 = bproj_point(g, w, x0, y0), a1 = bproj_point(g, w, x1, y1);
                double plo = std::min(a0, a1), phi = std::max(a0, a1), sp = phi - plo;
                double xcn = 0.5 * (x0 + x1), ycn = 0.5 * (y0 + y1);
                double wt = 1.0;
                if (g.scan == ST_ED) {
                    // magnification weight R^2/U^2, U = distance from source along the central ray
                    double U = g.s2r + (w.sn * xcn - w.cs * ycn);
                    wt = g.s2r * g.s2r / (U * U);
                } else if (g.scan == ST_EA) {
                    wt = 1.0 / ((xcn - w.sx) * (xcn - w.sx) + (ycn - w.sy) * (ycn - w.sy));
                }
                for (int j = 0; j < D; j++) {
                    double ov = std::min(hi[j], phi) - std::max(lo[j], plo);
                    if (!(ov > 0)) continue;
                    double fprj = g.scan == ST_PARA ? g.dImg / g.dDet : g.dImg / (cb[j] * (hi[j] - lo[j]));
                    for (int b = 0; b < B; b++) {
                        if (op == OP_PRJ) {
                            out[((size_t)b * V + v) * D + j] += ov * fprj * in[((size_t)b * H + r) * W + c];
                        } else if (op == OP_PRJ_T) {
                            out[((size_t)b * H + r) * W + c] += ov * fprj * in[((size_t)b * V + v) * D + j];
                        } else if (op == OP_BPRJ) {
                            out[((size_t)b * H + r) * W + c] += ov / sp * wt * in[((size_t)b * V + v) * D + j];
                        } else if (op == OP_BPRJ_SV) {
                            out[(((size_t)b * V + v) * H + r) * W + c] += ov / sp * wt * in[((size_t)b * V + v) * D + j];
                        } else {
                            double f = ov / sp * wt;
                            if (op == OP_BPRJ_T && g.scan == ST_EA) f *= g.dImg / cb[j];
                            out[((size_t)b * V + v) * D + j] += f * in[((size_t)b * H + r) * W + c];
                        }
                    }
                }
            }
        }
    }
}

}  // namespace

extern "C" {

int ctor_parse(const double* opt, int nopt, int* ints5, double* dbl7) {
    Geom g;
    int rc = parse(opt, nopt, &g);
    if (rc) return rc;
    ints5[0] = g.V; ints5[1] = g.D; ints5[2] = g.W; ints5[3] = g.H; ints5[4] = g.scan;
    dbl7[0] = g.dImg; dbl7[1] = g.dDet; dbl7[2] = g.Ang0; dbl7[3] = g.dAng;
    dbl7[4] = g.s2r; dbl7[5] = g.d2r; dbl7[6] = g.bs;
    return 0;
}

// Per-view plan as the oracle sees it: mode (1 = X), rev, and the D+1 internal edges.
int ctor_view(const double* opt, int nopt, int v, int* mode_rev, double* edges, double* cbin) {
    Geom g;
    int rc = parse(opt, nopt, &g);
    if (rc) return rc;
    if (v < 0 || v >= g.V) return E_ARG;
    View w;
    view_setup(g, v, w);
    mode_rev[0] = w.xmode; mode_rev[1] = w.rev;
    std::memcpy(edges, w.e.data(), sizeof(double) * (g.D + 1));
    if (cbin) std::memcpy(cbin, w.cb.data(), sizeof(double) * g.D);
    return 0;
}

// Apply one operator.  `chunk` <= 0 uses the reference's 256-bin chunks.
int ctor_apply(int op, const double* opt, int nopt, const double* in, double* out, int B, int chunk,
               int nthreads) {
    Geom g;
    int rc = parse(opt, nopt, &g);
    if (rc) return rc;
    if (B < 0 || !in || !out) return E_ARG;
    if (chunk <= 0) chunk = CHUNK_DEFAULT;
    set_threads(nthreads);
    std::vector<View> views = all_views(g);
    switch (op) {
    case OP_PRJ: run_to_sino(g, views, in, out, B, chunk, OP_PRJ); break;
    case OP_BPRJ_T: run_to_sino(g, views, in, out, B, chunk, OP_BPRJ_T); break;
    case OP_BPRJ_T_TRUE: run_to_sino(g, views, in, out, B, chunk, OP_BPRJ_T_TRUE); break;
    case OP_PRJ_T: run_to_image(g, views, in, out, B, chunk, OP_PRJ_T); break;
    case OP_BPRJ: run_to_image(g, views, in, out, B, chunk, OP_BPRJ); break;
    case OP_BPRJ_SV: run_sv(g, views, in, out, B, chunk); break;
    case OP_FBP: {
        std::vector<double> q((size_t)B * g.V * g.D);
        ramp_rows(g, in, q.data(), B * g.V);
        run_to_image(g, views, q.data(), out, B, chunk, OP_BPRJ);
        const double s = g.dAng / 2;   // fan_ed_kernel.cu:1380
        for (size_t i = 0; i < (size_t)B * g.H * g.W; i++) out[i] *= s;
        break;
    }
    default: return E_ARG;
    }
    return 0;
}

// Pre-weighted, ramp-filtered rows (the conv2d output of fbp) for `rows` sinogram rows.
int ctor_ramp(const double* opt, int nopt, const double* y, double* q, int rows, int nthreads) {
    Geom g;
    int rc = parse(opt, nopt, &g);
    if (rc) return rc;
    set_threads(nthreads);
    ramp_rows(g, y, q, rows);
    return 0;
}

int ctor_rl_taps(int D, double delta, double* h) {
    if (D < 1) return E_ARG;
    rl_taps(D, delta, h);
    return 0;
}

int ctor_fbp_preweight(const double* opt, int nopt, double* pw, double* delta) {
    Geom g;
    int rc = parse(opt, nopt, &g);
    if (rc) return rc;
    fbp_preweight(g, pw, delta);
    return 0;
}

int ctor_brute(int op, const double* opt, int nopt, const double* in, double* out, int B) {
    Geom g;
    int rc = parse(opt, nopt, &g);
    if (rc) return rc;
    if (op < 0 || op > 6 || op == OP_FBP) return E_ARG;
    brute(g, op, in, out, B);
    return 0;
}

int ctor_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
