// Host-side FP64 geometry plan (see plan.h).  Compiled by the host compiler
// with -ffp-contract=off so that the per-view axis decisions at pi/4 ties are
// reproducible bit for bit.
#include "plan.h"

#include <cmath>
#include <cstdio>

#include "../../include/ctlib_abi.h"

namespace ctlib {

namespace {

inline double map_x(double sx, double sy, double px, double py) { return (sx * py - sy * px) / (py - sy); }
inline double map_y(double sx, double sy, double px, double py) { return (sy * px - sx * py) / (px - sx); }

std::string fmt(const char* f, double a = 0, double b = 0) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), f, a, b);
    return buf;
}

// Odd polynomial for the fan-EA axis -> bin lookup: atan(t) on |t| <= T with
// T = tan(largest |fan angle| of a bin edge + one bin), from the Chebyshev
// expansion of atan(T x) on [-1, 1] (odd terms T_1 .. T_11, coefficients by the
// discrete cosine sum over 64 Chebyshev nodes), converted to powers of t.  A
// ray outside the fan has |t| > T and is clamped to T, which maps past the last
// bin edge as the exact arctangent would.  Kept only if the fp32-rounded
// coefficients stay within 5e-8 rad of atan on [0, T] (fans up to about 60
// degrees each side).
void fit_atan(const Geom& g, HostPlan* P) {
    const double PI = std::acos(-1.0);
    const double hD = g.dets / 2.0;
    const double gmax = std::fmax(std::fabs(-hD * g.dDet + g.binshift), std::fabs(hD * g.dDet + g.binshift)) + g.dDet;
    P->ea_poly = 0;
    if (!(gmax < 1.05)) return;
    const double T = std::tan(gmax);
    constexpr int N = 64, NT = 12;  // nodes; Chebyshev terms 0..11
    double a[NT] = {};
    for (int j = 1; j < NT; j += 2) {
        double s = 0.0;
        for (int k = 0; k < N; k++) {
            const double th = PI * (k + 0.5) / N;
            s += std::atan(T * std::cos(th)) * std::cos(j * th);
        }
        a[j] = 2.0 * s / N;
    }
    // power coefficients of sum_j a_j T_j(x) through T_{j+1} = 2x T_j - T_{j-1}
    double Tm[NT][NT] = {}, b[NT] = {};
    Tm[0][0] = 1.0;
    Tm[1][1] = 1.0;
    for (int j = 2; j < NT; j++)
        for (int i = 0; i < NT; i++) Tm[j][i] = (i > 0 ? 2.0 * Tm[j - 1][i - 1] : 0.0) - Tm[j - 2][i];
    for (int j = 1; j < NT; j += 2)
        for (int i = 0; i < NT; i++) b[i] += a[j] * Tm[j][i];
    // atan(t) ~ sum_k c_k t^(2k+1), c_k = b_(2k+1) / T^(2k+1)
    for (int k = 0; k < 6; k++) P->ea_c[k] = (float)(b[2 * k + 1] / std::pow(T, 2 * k + 1));
    P->ea_T = (float)T;
    double err = 0.0;
    for (int i = 0; i <= 4000; i++) {
        const double t = T * i / 4000.0, s = t * t;
        double p = P->ea_c[5];
        for (int k = 4; k >= 0; k--) p = p * s + P->ea_c[k];
        err = std::fmax(err, std::fabs(t * p - std::atan(t)));
    }
    P->ea_poly = err <= 5e-8 ? 1 : 0;
}

}  // namespace

int parse_options(const double* o, int n, Geom* g, std::string* why) {
    if (!o || n < 1) { *why = "options vector is empty"; return CTLIB_ERR_INVALID_ARG; }
    // scan_type is the last entry (ctlib.cpp:18); the reference exit(0)s on anything else.
    double st = o[n - 1];
    if (!(st == 0.0 || st == 1.0 || st == 2.0)) {
        *why = fmt("scan_type (last options entry) must be 0, 1 or 2, got %g", st);
        return CTLIB_ERR_BAD_SCAN_TYPE;
    }
    g->scan_type = (int)st;
    // Geometry entries: fan uses indices 0..10 (fan_ed_kernel.cu:1165-1175); parallel uses
    // 0..8 with binshift = options[8] (para_kernel.cu:687-695).
    const int need = g->scan_type == SCAN_PARA ? 9 : 11;
    if (n < need) {
        *why = fmt(g->scan_type == SCAN_PARA ? "parallel-beam options need >= 9 entries, got %g"
                                             : "fan-beam options need >= 11 entries, got %g",
                   (double)n);
        return CTLIB_ERR_INVALID_ARG;
    }
    static const char* names[4] = {"views", "dets", "width", "height"};
    for (int i = 0; i < 4; i++) {
        if (!(o[i] >= 1.0) || o[i] != std::floor(o[i]) || o[i] > (1 << 20)) {
            *why = std::string(names[i]) + fmt(" must be a positive integer, got %g", o[i]);
            return CTLIB_ERR_INVALID_ARG;
        }
    }
    g->views = (int)o[0];
    g->dets = (int)o[1];
    g->width = (int)o[2];
    g->height = (int)o[3];
    g->dImg = o[4];
    g->dDet = o[5];
    g->Ang0 = o[6];
    g->dAng = o[7];
    if (g->scan_type == SCAN_PARA) {
        g->s2r = g->d2r = 0.0;
        g->binshift = o[8];
    } else {
        g->s2r = o[8];
        g->d2r = o[9];
        g->binshift = o[10];
    }
    if (!(g->dImg > 0) || !std::isfinite(g->dImg)) { *why = "dImg must be > 0"; return CTLIB_ERR_INVALID_ARG; }
    if (!(g->dDet > 0) || !std::isfinite(g->dDet)) { *why = "dDet must be > 0"; return CTLIB_ERR_INVALID_ARG; }
    if (!std::isfinite(g->Ang0) || !std::isfinite(g->dAng) || !std::isfinite(g->binshift)) {
        *why = "Ang0, dAng and binshift must be finite";
        return CTLIB_ERR_INVALID_ARG;
    }
    if (g->scan_type != SCAN_PARA && (!(g->s2r > 0) || !std::isfinite(g->s2r) || !std::isfinite(g->d2r) ||
                                      !(g->s2r + g->d2r > 0))) {
        *why = "fan beam needs s2r > 0 and s2r + d2r > 0";
        return CTLIB_ERR_INVALID_ARG;
    }
    return CTLIB_OK;
}

int build_plan(const Geom& g, HostPlan* P, std::string* why) {
    const int V = g.views, D = g.dets, W = g.width, H = g.height;
    const double PI = std::acos(-1.0);
    const double hD = (double)D / 2;
    const double R = g.s2r;
    const double mag = g.scan_type == SCAN_FAN_ED ? g.s2r / (g.s2r + g.d2r) : 1.0;
    const double vd = mag * g.dDet;      // virtual bin width (fan_ed_kernel.cu:68)
    const double vs = mag * g.binshift;  // virtual shift (fan_ed_kernel.cu:69)

    P->g = g;
    P->max_lines = H > W ? H : W;
    P->vp.assign(V, ViewParams{});
    P->edges.assign((size_t)V * (D + 1), 0.0);
    P->post_prj.assign((size_t)V * D, 0.0);
    P->post_bprj_t.assign((size_t)V * D, 1.0);
    P->lines.assign((size_t)V * P->max_lines, LineCoef{});
    P->adj_lk.assign((size_t)V * P->max_lines * 4, 0.0f);
    P->adj_wk.assign((size_t)V * P->max_lines * 4, 0.0f);
    P->lookup_d.assign((size_t)V * 6, 0.0);
    P->xviews.clear();
    P->yviews.clear();

    for (int v = 0; v < V; v++) {
        // Per-view setup of fan_ed_kernel.cu:60-77 with the angle formed in double.
        const double ang = (double)v * g.dAng + g.Ang0;
        const double err = std::fabs(ang - std::round(ang / PI) * PI) * 4 / PI;
        const bool xmode = err <= 1;
        const double err2 = (ang - std::floor(ang / 2 / PI) * 2 * PI) * 4 / PI;
        const bool rev = err2 >= 3 && err2 < 7;
        const double cs = std::cos(ang), sn = std::sin(ang);
        const double sx = -sn * R, sy = cs * R;
        const double sig = rev ? -1.0 : 1.0;

        ViewParams& vp = P->vp[v];
        vp.sx = sx;
        vp.sy = sy;
        vp.xmode = xmode;
        vp.rev = rev;
        vp.nlines = xmode ? H : W;
        vp.npix = xmode ? W : H;
        (xmode ? P->xviews : P->yviews).push_back(v);

        // Bin edges on the sweep axis, internal order (fan_ed_kernel.cu:78-96;
        // fan_ea_kernel.cu:70-84; para_kernel.cu:50-60).
        double* e = &P->edges[(size_t)v * (D + 1)];
        for (int i = 0; i <= D; i++) {
            double ev;
            if (g.scan_type == SCAN_FAN_ED) {
                const double u = rev ? (hD - i) * vd + vs : (i - hD) * vd + vs;
                const double px = u * cs, py = u * sn;
                ev = xmode ? map_x(sx, sy, px, py) : map_y(sx, sy, px, py);
            } else if (g.scan_type == SCAN_FAN_EA) {
                const double beta = rev ? (hD - i) * g.dDet + g.binshift + ang : (i - hD) * g.dDet + g.binshift + ang;
                ev = xmode ? sx + std::sin(beta) / std::cos(beta) * sy : sy + std::cos(beta) / std::sin(beta) * sx;
            } else {
                const double t = rev ? (hD - i) * g.dDet + g.binshift : (i - hD) * g.dDet + g.binshift;
                ev = xmode ? t / cs : t / sn;
            }
            if (!std::isfinite(ev)) {
                *why = fmt("view %g: bin edge %g is not finite", v, i);
                return CTLIB_ERR_GEOMETRY;
            }
            e[i] = ev;
        }
        for (int i = 0; i < D; i++) {
            if (!(e[i + 1] > e[i])) {
                *why = fmt("view %g: detector edges are not increasing on the sweep axis at bin %g", v, i);
                return CTLIB_ERR_GEOMETRY;
            }
        }

        // Bin factors: c_k from fan_ed_kernel.cu:146-154 (X) / :235-243 (Y); w_k = e_{k+1} - e_k.
        double* pp = &P->post_prj[(size_t)v * D];
        double* pb = &P->post_bprj_t[(size_t)v * D];
        auto kap = [&](double ev) {
            return xmode ? std::fabs(sy) / std::sqrt((ev - sx) * (ev - sx) + sy * sy)
                         : std::fabs(sx) / std::sqrt((ev - sy) * (ev - sy) + sx * sx);
        };
        for (int k = 0; k < D; k++) {
            if (g.scan_type == SCAN_PARA) {
                pp[k] = g.dImg / g.dDet;  // dinterval, para_kernel.cu:43
            } else {
                const double c = (kap(e[k]) + kap(e[k + 1])) / 2;
                pp[k] = g.dImg / c / (e[k + 1] - e[k]);
                if (g.scan_type == SCAN_FAN_EA) pb[k] = g.dImg / c;  // fan_ea_kernel.cu:590-607
            }
        }

        // Axis coordinate -> (internal bin index + 1): fp64 coefficients for fp64 data,
        // rounded to fp32 in the view record for fp32 data.
        double* lkd = &P->lookup_d[(size_t)v * 6];  // lk0..lk3, ea_sc, ea_off
        if (g.scan_type == SCAN_FAN_ED) {
            const double K = hD + 1 - sig * vs / vd;
            if (xmode) {
                lkd[0] = sig * R * cs / vd + K * sn;
                lkd[1] = K * R;
                lkd[2] = sn;
                lkd[3] = R;
            } else {
                lkd[0] = sig * R * sn / vd - K * cs;
                lkd[1] = K * R;
                lkd[2] = -cs;
                lkd[3] = R;
            }
        } else if (g.scan_type == SCAN_FAN_EA) {
            lkd[0] = xmode ? cs : sn;
            lkd[1] = 0.0;
            lkd[2] = xmode ? sn : -cs;
            lkd[3] = R;
            lkd[4] = sig / g.dDet;
            lkd[5] = hD + 1 - sig * g.binshift / g.dDet;
        } else {
            lkd[0] = sig * (xmode ? cs : sn) / g.dDet;
            lkd[1] = hD + 1 - sig * g.binshift / g.dDet;
            lkd[2] = 0.0;
            lkd[3] = 1.0;
        }
        for (int k = 0; k < 4; k++) vp.lk[k] = (float)lkd[k];
        vp.ea_sc = (float)lkd[4];
        vp.ea_off = (float)lkd[5];

        // Per-line affine pixel-edge maps and pixel weights.
        LineCoef* lc = &P->lines[(size_t)v * P->max_lines];
        const int nl = vp.nlines;
        const double x0 = -W / 2.0 * g.dImg, y0 = -H / 2.0 * g.dImg;
        for (int l = 0; l < nl; l++) {
            LineCoef c{};
            if (xmode) {
                const double yl = (H / 2.0 - l - 0.5) * g.dImg;  // fan_ed_kernel.cu:106
                if (g.scan_type == SCAN_PARA) {
                    c.a = sn / cs * yl + x0;  // para_kernel.cu:66
                    c.s = g.dImg;
                } else {
                    const double den = yl - sy;
                    c.a = (sx * yl - sy * x0) / den;
                    c.s = -sy * g.dImg / den;
                }
                if (g.scan_type == SCAN_FAN_ED) {
                    c.w0 = 1 - (sx * (x0 + 0.5 * g.dImg) + sy * yl) / (R * R);
                    c.w1 = -sx * g.dImg / (R * R);
                    c.w2 = 0;
                } else if (g.scan_type == SCAN_FAN_EA) {
                    c.w0 = sx - x0 - 0.5 * g.dImg;
                    c.w1 = -g.dImg;
                    c.w2 = (sy - yl) * (sy - yl);
                } else {
                    c.w0 = 1; c.w1 = 0; c.w2 = 0;
                }
            } else {
                const double xl = (l - W / 2.0 + 0.5) * g.dImg;  // fan_ed_kernel.cu:195
                if (g.scan_type == SCAN_PARA) {
                    c.a = y0 + cs / sn * xl;  // para_kernel.cu:126
                    c.s = g.dImg;
                } else {
                    const double den = xl - sx;
                    c.a = (sy * xl - sx * y0) / den;
                    c.s = -sx * g.dImg / den;
                }
                if (g.scan_type == SCAN_FAN_ED) {
                    c.w0 = 1 - (sx * xl + sy * (y0 + 0.5 * g.dImg)) / (R * R);
                    c.w1 = -sy * g.dImg / (R * R);
                    c.w2 = 0;
                } else if (g.scan_type == SCAN_FAN_EA) {
                    c.w0 = sy - y0 - 0.5 * g.dImg;
                    c.w1 = -g.dImg;
                    c.w2 = (sx - xl) * (sx - xl);
                } else {
                    c.w0 = 1; c.w1 = 0; c.w2 = 0;
                }
            }
            if (!(c.s > 0) || !std::isfinite(c.a)) {
                *why = fmt("view %g line %g: pixel edges do not increase along the sweep axis "
                           "(source inside the image?)", v, l);
                return CTLIB_ERR_GEOMETRY;
            }
            c.A = 1.0 / c.s;
            c.B = 1.0 - c.a / c.s;
            c.inv_s = 1.0 / c.s;
            lc[l] = c;
            // the view's lookup (lkd) at pixel edge p of this line, lk(a + s p) = (c0 p + c1) / (c2 p + c3),
            // composed in fp64; c0 and c2 are stored times 64 (the kernels step p in 64s)
            float* ck = &P->adj_lk[((size_t)v * P->max_lines + l) * 4];
            ck[0] = (float)(64.0 * lkd[0] * c.s);
            ck[1] = (float)(lkd[0] * c.a + lkd[1]);
            ck[2] = (float)(64.0 * lkd[2] * c.s);
            ck[3] = (float)(lkd[2] * c.a + lkd[3]);
            const double cw = std::sqrt(c.s);
            float* wk = &P->adj_wk[((size_t)v * P->max_lines + l) * 4];
            wk[0] = (float)(64.0 * cw * c.w1);
            wk[1] = (float)(cw * c.w1);
            wk[2] = (float)(cw * c.w0);
            wk[3] = (float)(c.s * c.w2);
        }
    }

    if (g.scan_type == SCAN_FAN_EA) fit_atan(g, P);

    // fbp: RL taps (fan_ed_kernel.cu:925-943) and pre-weight (fan_ed_kernel.cu:1349-1355;
    // fan_ea_kernel.cu:1075-1080; para_kernel.cu:823), both in output bin order.
    const double delta = g.scan_type == SCAN_FAN_ED ? g.dDet * g.s2r / (g.s2r + g.d2r) : g.dDet;
    P->ramp_taps.assign(2 * (size_t)D - 1, 0.0);
    for (int t = 0; t < 2 * D - 1; t++) {
        const int x = t - D + 1;
        if (std::abs(x) % 2 == 1) P->ramp_taps[t] = -1.0 / (PI * PI * x * x * delta * delta);
        else if (x == 0) P->ramp_taps[t] = 1.0 / (4 * delta * delta);
    }
    P->ramp_preweight.assign(D, 0.0);
    for (int j = 0; j < D; j++) {
        const double r = -hD + 0.5 + j;
        if (g.scan_type == SCAN_FAN_ED) {
            const double t = r * delta;
            P->ramp_preweight[j] = g.s2r / std::sqrt(g.s2r * g.s2r + t * t) * delta;
        } else if (g.scan_type == SCAN_FAN_EA) {
            P->ramp_preweight[j] = std::cos(r * g.dDet) * g.s2r * g.dDet;
        } else {
            P->ramp_preweight[j] = g.dDet;
        }
    }
    return CTLIB_OK;
}

}  // namespace ctlib
