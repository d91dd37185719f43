// Host-side geometry plan for the distance-driven projector family.
//
// Everything here is computed once per geometry, in double precision, and is
// shared by every kernel launch that uses the same options vector.  The plan
// turns the reference's per-block setup (fan_ed_kernel.cu:58-102, and the twins
// in fan_ea_kernel.cu / para_kernel.cu) into tables:
//   * per view: sweep axis (X: project onto y=0 and sweep rows; Y: project onto
//     x=0 and sweep columns), reversal flag, source position;
//   * per view: the D+1 detector-bin edges mapped onto the sweep axis, in
//     increasing ("internal") order, and the bin normalisation factors;
//   * per (view, line): the affine map from pixel-edge index p to the axis
//     coordinate  pi(p) = a + s*p  (the closed form of the reference's
//     incremental i0x += dImage), and its inverse tau'(e) = A*e + B;
//   * per view: coefficients of the axis-coordinate -> bin-index lookup.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ctlib {

enum ScanType { SCAN_FAN_ED = 0, SCAN_FAN_EA = 1, SCAN_PARA = 2 };

struct Geom {
    int views, dets, width, height, scan_type;
    double dImg, dDet, Ang0, dAng, s2r, d2r, binshift;
};

// Per-view parameters read by the kernels (one 64-byte record per view).
struct alignas(16) ViewParams {
    double sx, sy;        // source position (fan); unused for parallel beam
    float lk[4];          // lookup: bin+1 = (lk0*e + lk1)/(lk2*e + lk3)            (ED, para)
                          //   EA: t = (lk0*e + lk1)/(lk2*e + lk3), bin+1 = atan(t)*ea_sc + ea_off
    float ea_sc, ea_off;
    int xmode, rev;       // sweep axis (1 = rows, projected onto y=0) and bin-order reversal
    int nlines, npix;     // lines swept by this view and pixels per line
};

// Per (view, line) coefficients.
//   pixel edge p sits at  a + s*p  on the sweep axis;
//   tau'(e) = A*e + B  is the inverse map, shifted by +1 so that the table index is trunc(tau');
//   backprojection pixel weight of pixel p on this line:  1 / (s * ((w0 + w1*p)^2 + w2))
//   (fan-ED R^2/U^2 with U/R = w0 + w1*p, w2 = 0; fan-EA 1/|src - pixel|^2; parallel 1/s).
struct alignas(64) LineCoef {
    double a, s, A, B, w0, w1, w2, inv_s;
};

struct HostPlan {
    Geom g;
    int max_lines;                       // max(H, W)
    std::vector<ViewParams> vp;          // [V]
    std::vector<double> edges;           // [V][D+1] internal order, increasing
    std::vector<double> post_prj;        // [V][D]   projection factor dImg/(c_k w_k) (fan) or dImg/dDet (para)
    std::vector<double> post_bprj_t;     // [V][D]   backprojection_t factor: 1, or dImg/c_k for fan-EA
    std::vector<LineCoef> lines;         // [V][max_lines]
    std::vector<double> lookup_d;        // [V][6] ViewParams lk0..lk3, ea_sc, ea_off in fp64
    // [V][max_lines][4] the view's lookup composed with the line's pixel-edge map, fp32:
    // {64*lk0*s, lk0*a + lk1, 64*lk2*s, lk2*a + lk3}, so that at pixel edge p the lookup
    // argument is (c0*p/64 + c1) / (c2*p/64 + c3) (the bin-to-pixel kernels' fp32 path)
    std::vector<float> adj_lk;
    // [V][max_lines][4] the line's pixel weight inv_s / ((w0 + w1 p)^2 + w2) rewritten for fp32:
    // with c = sqrt(s), {64 c w1, c w1, c w0, c^2 w2}, so that at pixel p = lane + 64 k the
    // weight is 1 / (u^2 + c^2 w2), u = k * (64 c w1) + (lane * c w1 + c w0) (the weighted
    // bin-to-pixel kernel's fp32 path; the call's scale multiplies its output)
    std::vector<float> adj_wk;
    std::vector<int> xviews, yviews;     // view indices per sweep axis, ascending
    std::vector<double> ramp_taps;       // [2D-1] RL taps h(-(D-1)..D-1)
    std::vector<double> ramp_preweight;  // [D] fbp pre-weight (output bin order)
    // fan-EA lookup: atan(t) ~ t * (c0 + c1 t^2 + ... + c5 t^10) for |t| <= ea_T (t clamped there),
    // used when ea_poly (the fit is within 5e-8 rad of atan on [0, ea_T]); else the kernels call atanf
    float ea_c[6] = {0, 0, 0, 0, 0, 0};
    float ea_T = 0.0f;
    int ea_poly = 0;
};

// Parses the reference options vector (ctlib.cpp:18; fan_ed_kernel.cu:1165-1175;
// para_kernel.cu:687-695).  Returns 0, or a ctlib_status code.
int parse_options(const double* opts, int n, Geom* out, std::string* why);

// Builds the full plan.  Returns 0, or a ctlib_status code with a reason.
int build_plan(const Geom& g, HostPlan* out, std::string* why);

}  // namespace ctlib
