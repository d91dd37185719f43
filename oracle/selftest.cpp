// Self-test driver for the oracle's sanitizer build (test infrastructure only).
// Built with -fsanitize=address,undefined together with ct_oracle.cpp; it walks
// every operator, chunk width and thread count on small geometries so that an
// out-of-bounds table access or undefined arithmetic in the oracle shows up as
// a sanitizer report.  Exit status 0 means every check below held.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
int ctor_parse(const double* opt, int nopt, int* ints5, double* dbl7);
int ctor_apply(int op, const double* opt, int nopt, const double* in, double* out, int B, int chunk,
               int nthreads);
int ctor_brute(int op, const double* opt, int nopt, const double* in, double* out, int B);
int ctor_ramp(const double* opt, int nopt, const double* y, double* q, int rows, int nthreads);
}

namespace {

int failures = 0;

void check(bool ok, const char* what, int scan, int op) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s (scan %d, op %d)\n", what, scan, op);
        failures++;
    }
}

double rel(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1));
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

// Option vectors in the reference's layout: views, dets, width, height, dImg, dDet,
// Ang0, dAng, s2r, d2r, binshift, scan_type (9 entries for parallel beams).
std::vector<double> geometry(int scan, int views, int dets, int w, int h, double ang0) {
    const double pi = std::acos(-1.0);
    if (scan == 2) return {double(views), double(dets), double(w), double(h), 0.1, 0.13, ang0, pi / views, 0.02, 2};
    double ddet = scan == 1 ? 0.03 : 0.13;
    return {double(views), double(dets), double(w), double(h), 0.1, ddet, ang0, 2 * pi / views, 3.0, 2.0, 0.02,
            double(scan)};
}

}  // namespace

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const int B = 2;
    for (int scan = 0; scan < 3; scan++) {
        for (double ang0 : {0.1, 0.7853981633974483}) {   // generic start, and a start on a 45-degree switch
            std::vector<double> o = geometry(scan, 20, 17, 12, 10, ang0);
            const int n = (int)o.size();
            int ints[5];
            double dbl[7];
            check(ctor_parse(o.data(), n, ints, dbl) == 0, "parse", scan, -1);
            const size_t img = (size_t)B * 12 * 10, sino = (size_t)B * 20 * 17;
            std::vector<double> x(img), y(sino);
            for (auto& v : x) v = u(rng);
            for (auto& v : y) v = u(rng);
            for (int op = 0; op <= 6; op++) {
                const bool from_img = op == 0 || op == 3 || op == 4;
                const std::vector<double>& in = from_img ? x : y;
                size_t nout = (op == 0 || op == 3 || op == 4) ? sino : (op == 6 ? (size_t)B * 20 * 12 * 10 : img);
                std::vector<double> a(nout), b(nout), c(nout);
                check(ctor_apply(op, o.data(), n, in.data(), a.data(), B, 0, 1) == 0, "apply", scan, op);
                check(ctor_apply(op, o.data(), n, in.data(), b.data(), B, 3, 3) == 0, "apply chunk 3", scan, op);
                check(rel(b, a) < 1e-12, "chunk width / threads change the result", scan, op);
                if (op != 5) {
                    check(ctor_brute(op, o.data(), n, in.data(), c.data(), B) == 0, "brute", scan, op);
                    check(rel(a, c) < 1e-10, "merge differs from brute force", scan, op);
                }
            }
            // <A x, y> = <x, A^T y> for the projection pair.
            std::vector<double> ax(sino), aty(img);
            ctor_apply(0, o.data(), n, x.data(), ax.data(), B, 0, 2);
            ctor_apply(1, o.data(), n, y.data(), aty.data(), B, 0, 2);
            double l = dot(ax, y), r = dot(x, aty);
            check(std::fabs(l - r) <= 1e-12 * std::fabs(l), "projection adjoint", scan, 0);
            std::vector<double> q(sino);
            check(ctor_ramp(o.data(), n, y.data(), q.data(), B * 20, 2) == 0, "ramp", scan, 5);
            for (double v : q) check(std::isfinite(v), "ramp finite", scan, 5);
        }
    }
    // Rejected options: unknown scan type, short vector, non-integer sizes, zero pixel size.
    std::vector<double> bad = geometry(0, 20, 17, 12, 10, 0.1);
    int ints[5];
    double dbl[7];
    bad.back() = 3;
    check(ctor_parse(bad.data(), (int)bad.size(), ints, dbl) == 2, "scan type 3 accepted", 3, -1);
    bad.back() = 0;
    std::vector<double> shortv(bad.begin(), bad.begin() + 9);   // fan vector missing d2r and binshift
    shortv.push_back(0);
    check(ctor_parse(shortv.data(), (int)shortv.size(), ints, dbl) == 1, "short vector accepted", 0, -1);
    bad[1] = 16.5;
    check(ctor_parse(bad.data(), (int)bad.size(), ints, dbl) == 1, "fractional dets accepted", 0, -1);
    bad[1] = 17;
    bad[4] = 0;
    check(ctor_parse(bad.data(), (int)bad.size(), ints, dbl) == 1, "zero pixel size accepted", 0, -1);
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("oracle selftest ok\n");
    return 0;
}
