// CPU check of the periodic-tridiagonal Givens QR inverse
// (fdfd.jl_amd/csrc/tri_qr.h): for random complex blocks (indefinite
// diagonals included) of every small size and a few large ones, the columns
// formed exactly as the device kernel forms them must give T X = I to
// round-off, and the solve of T x = v for a random v through apply_q_vec
// must leave a residual at round-off.  Prints one line per case; exit
// status 1 on failure.
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "tri_qr.h"

using namespace fdfd::tri;
using cd = std::complex<double>;

static double run(int n, double diag_scale, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<cz> d(n), e(n), f(n);
    for (int i = 0; i < n; ++i) {
        d[i] = mkz(diag_scale * U(rng), diag_scale * U(rng));
        e[i] = mkz(U(rng), U(rng));
        f[i] = mkz(U(rng), U(rng));
    }
    std::vector<double> w(work_doubles(n));
    double pmin, pmax;
    auto get = [&](int i, cz& dd, cz& ee, cz& ff) { dd = d[i]; ee = e[i]; ff = f[i]; };
    if (factor(n, get, w.data(), pmin, pmax) != 0) return 1e300;
    std::vector<cd> X((size_t)n * n);
    for (int b = 0; b < n; ++b) {
        std::vector<cz> col(n, mkz(0, 0));
        apply_q(n, w.data(), b, 0, [&](int k, cz v) { col[k] = v; });
        apply_rinv(n, w.data(), b, [&](int k) { return col[k]; },
                   [&](int k, cz v) { col[k] = v; });
        for (int k = 0; k < n; ++k) X[(size_t)k * n + b] = cd(col[k].re, col[k].im);
    }
    // T X - I, with T assembled densely (wrap entries merged)
    std::vector<cd> T((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        T[(size_t)i * n + i] += cd(d[i].re, d[i].im);
        T[(size_t)i * n + (i + 1) % n] += cd(e[i].re, e[i].im);
        T[(size_t)i * n + (i + n - 1) % n] += cd(f[i].re, f[i].im);
    }
    double err = 0, xn = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            cd s = 0;
            for (int k = 0; k < n; ++k) s += T[(size_t)i * n + k] * X[(size_t)k * n + j];
            err = std::max(err, std::abs(s - (i == j ? 1.0 : 0.0)));
            xn = std::max(xn, std::abs(X[(size_t)i * n + j]));
        }
    // general right-hand side through apply_q_vec (in place) + apply_rinv:
    // T x = v to round-off
    std::vector<cz> v(n), x(n);
    for (int i = 0; i < n; ++i) v[i] = mkz(U(rng), U(rng));
    x = v;
    apply_q_vec(n, w.data(), [&](int k) { return x[k]; }, [&](int k, cz y) { x[k] = y; });
    apply_rinv(n, w.data(), 0, [&](int k) { return x[k]; }, [&](int k, cz y) { x[k] = y; });
    double verr = 0, vn = 0;
    for (int i = 0; i < n; ++i) {
        cd s = 0;
        for (int k = 0; k < n; ++k) s += T[(size_t)i * n + k] * cd(x[k].re, x[k].im);
        verr = std::max(verr, std::abs(s - cd(v[i].re, v[i].im)));
        vn = std::max(vn, std::abs(cd(v[i].re, v[i].im)));
    }
    return std::max(err / std::max(1.0, xn), verr / (vn * std::max(1.0, xn)));
}

int main() {
    int bad = 0;
    unsigned seed = 11;
    const int ns[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 17, 33, 64, 100, 257};
    for (int n : ns)
        for (double ds : {4.0, 0.3}) {
            const double r = run(n, ds, seed++);
            const bool ok = r < 1e-12;
            bad += !ok;
            std::printf("n=%d diag=%.1f |TX-I|/max|X|=%.3e %s\n", n, ds, r, ok ? "ok" : "FAIL");
        }
    return bad ? 1 : 0;
}
