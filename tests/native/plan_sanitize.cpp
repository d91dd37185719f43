// Host-plan self-test for the sanitizer build (test infrastructure only).
// Compiled with -fsanitize=address,undefined together with the product's
// plan.cpp; it parses and plans a spread of geometries (minimal sizes, the
// 45-degree axis switch, wide and narrow fan-EA fans, source-inside-image and
// rejected vectors) so that any out-of-bounds table write or undefined
// arithmetic in the host planner shows up as a sanitizer report.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../ctlib_amd/csrc/plan.h"

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s\n", what.c_str());
        failures++;
    }
}

bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

// Returns the build_plan status (or the parse status when parsing fails).
int plan(const std::vector<double>& o, const std::string& name) {
    ctlib::Geom g;
    ctlib::HostPlan P;
    std::string why;
    int rc = ctlib::parse_options(o.data(), (int)o.size(), &g, &why);
    if (rc) return rc;
    rc = ctlib::build_plan(g, &P, &why);
    if (rc) {
        std::fprintf(stderr, "%s: status %d: %s\n", name.c_str(), rc, why.c_str());
        return rc;
    }
    const size_t V = g.views, D = g.dets;
    expect(P.vp.size() == V, name + ": view records");
    expect(P.edges.size() == V * (D + 1), name + ": edge table size");
    expect(P.post_prj.size() == V * D && P.post_bprj_t.size() == V * D, name + ": factor tables");
    expect(P.lines.size() == V * (size_t)P.max_lines, name + ": line table size");
    expect(P.xviews.size() + P.yviews.size() == V, name + ": axis split");
    expect(P.ramp_taps.size() == 2 * D - 1 && P.ramp_preweight.size() == D, name + ": ramp tables");
    expect(finite(P.edges) && finite(P.post_prj) && finite(P.post_bprj_t) && finite(P.ramp_taps) &&
               finite(P.ramp_preweight), name + ": non-finite table entry");
    for (size_t v = 0; v < V; v++)
        for (size_t i = 0; i < D; i++)
            expect(P.edges[v * (D + 1) + i] < P.edges[v * (D + 1) + i + 1], name + ": edges not increasing");
    return rc;
}

std::vector<double> fan(int V, int D, int W, int H, double dImg, double dDet, double ang0, double dAng, double s2r,
                        double d2r, double bs, int scan) {
    return {double(V), double(D), double(W), double(H), dImg, dDet, ang0, dAng, s2r, d2r, bs, double(scan)};
}

std::vector<double> para(int V, int D, int W, int H, double dImg, double dDet, double ang0, double dAng, double bs) {
    return {double(V), double(D), double(W), double(H), dImg, dDet, ang0, dAng, bs, 2};
}

}  // namespace

int main() {
    const double pi = std::acos(-1.0);
    // Accepted geometries.
    for (int scan = 0; scan < 2; scan++) {
        double dd = scan == 1 ? 0.0011 : 0.0032;
        expect(plan(fan(64, 736, 256, 256, 0.006641, dd, 0, 2 * pi / 64, 5.95, 4.906, 0, scan), "clinical") == 0,
               "clinical fan plan");
        expect(plan(fan(1, 1, 1, 1, 0.1, 0.13, 0.3, 0.1, 3, 2, 0, scan), "minimal") == 0, "minimal fan plan");
        expect(plan(fan(8, 17, 12, 10, 0.1, dd * 40, pi / 4, pi / 4, 3, 2, 0.02, scan), "switch") == 0,
               "45-degree switch plan");
        // binshift is a length for fan-ED and an angle (rad) for fan-EA
        const double bs = scan == 1 ? -0.05 : -0.7, ddw = scan == 1 ? 0.004 : 0.02;
        expect(plan(fan(36, 201, 64, 48, 0.02, ddw, -1.3, -2 * pi / 36, 2.0, 1.5, bs, scan),
                    "wide fan, reversed rotation") == 0, "wide fan plan");
    }
    expect(plan(para(180, 367, 256, 256, 0.0039, 0.0039, 0, pi / 180, 0), "para") == 0, "parallel plan");
    expect(plan(para(1, 1, 1, 1, 1, 1, 0, 0, 0), "para minimal") == 0, "minimal parallel plan");
    // A source inside the image cannot be swept in the DD order: a geometry error, not a crash.
    int rc = plan(fan(16, 33, 64, 64, 0.2, 0.1, 0.1, 2 * pi / 16, 3.0, 2.0, 0, 0), "source inside");
    expect(rc == 7, "source inside the image gives a geometry error");
    // Rejected vectors.
    expect(plan({}, "empty") == 1, "empty vector");
    expect(plan({8, 17, 12, 10, 0.1, 3}, "scan 3") == 2, "scan type 3");
    expect(plan(fan(8, 17, 12, 10, 0.1, 0.1, 0, 0.1, 0, 2, 0, 0), "s2r 0") == 1, "zero s2r");
    expect(plan(fan(8, 17, 12, 10, 0.1, 0.1, NAN, 0.1, 3, 2, 0, 0), "nan") == 1, "NaN angle");
    expect(plan(fan(8, 17, 0, 10, 0.1, 0.1, 0, 0.1, 3, 2, 0, 0), "width 0") == 1, "zero width");
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("plan selftest ok\n");
    return 0;
}
