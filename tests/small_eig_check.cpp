// CPU check of the host/device small dense eigensolver (fdfd.jl_amd/csrc/
// small_eig.h), the code that both Arnoldi drivers run on their projected
// matrices.  Two checks per case, each failing for a plausible slip:
//   * eigenpair residual  max_k |A v_k - l_k v_k| / max|A| <= 1e-12 on random
//     complex matrices (general and already-Hessenberg input);
//   * a known spectrum: A = U diag(l) U^H with U unitary (Gram-Schmidt of a
//     random matrix) and prescribed l, including a repeated value and values
//     of very different magnitude; every prescribed l must be found to
//     1e-11 relative.
// Prints one line per case; exit status 1 on failure.
#include <algorithm>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "small_eig.h"

using cd = std::complex<double>;
using fdfd::eig::cz;

static bool run(int n, const std::vector<cd>& A, std::vector<cd>& ev, std::vector<cd>& V) {
    std::vector<cz> a(n * n), q(n * n), e(n), vv(n * n), z(n), v(n);
    for (int i = 0; i < n * n; ++i) a[i] = fdfd::eig::mkz(A[i].real(), A[i].imag());
    if (fdfd::eig::small_eig(n, a.data(), q.data(), e.data(), vv.data(), z.data(), v.data()))
        return false;
    ev.resize(n);
    V.resize(n * n);
    for (int i = 0; i < n; ++i) ev[i] = cd(e[i].re, e[i].im);
    for (int i = 0; i < n * n; ++i) V[i] = cd(vv[i].re, vv[i].im);
    return true;
}

static double residual(int n, const std::vector<cd>& A, const std::vector<cd>& ev,
                       const std::vector<cd>& V) {
    double anorm = 0, worst = 0;
    for (auto& x : A) anorm = std::max(anorm, std::abs(x));
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            cd s = 0;
            for (int j = 0; j < n; ++j) s += A[i * n + j] * V[j * n + k];
            worst = std::max(worst, std::abs(s - ev[k] * V[i * n + k]) / anorm);
        }
    return worst;
}

int main() {
    std::mt19937_64 g(7);
    std::normal_distribution<double> nd;
    int bad = 0;
    for (int n : {1, 2, 3, 5, 8, 13, 20, 33, 41, 64}) {
        for (int rep = 0; rep < 2; ++rep) {
            std::vector<cd> A(n * n), ev, V;
            for (auto& x : A) x = cd(nd(g), nd(g));
            if (rep == 1)
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j + 1 < i; ++j) A[i * n + j] = 0;
            const bool ok = run(n, A, ev, V);
            const double r = ok ? residual(n, A, ev, V) : 1e300;
            const bool pass = r <= 1e-12;
            bad += !pass;
            std::printf("random n=%d %s residual=%.2e %s\n", n, rep ? "hessenberg" : "general", r,
                        pass ? "ok" : "FAIL");
        }
        // prescribed spectrum
        std::vector<cd> U(n * n), lam(n);
        for (auto& x : U) x = cd(nd(g), nd(g));
        for (int k = 0; k < n; ++k) {  // columns orthonormal
            for (int pass = 0; pass < 2; ++pass)
                for (int p = 0; p < k; ++p) {
                    cd s = 0;
                    for (int i = 0; i < n; ++i) s += std::conj(U[i * n + p]) * U[i * n + k];
                    for (int i = 0; i < n; ++i) U[i * n + k] -= s * U[i * n + p];
                }
            double nk = 0;
            for (int i = 0; i < n; ++i) nk += std::norm(U[i * n + k]);
            for (int i = 0; i < n; ++i) U[i * n + k] /= std::sqrt(nk);
        }
        for (int k = 0; k < n; ++k) lam[k] = cd(1.0 + k, 0.5 * k - 2.0) * (k % 3 == 0 ? 1e-3 : 1.0);
        if (n >= 3) lam[2] = lam[1];
        std::vector<cd> A(n * n, 0.0), ev, V;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) A[i * n + j] += U[i * n + k] * lam[k] * std::conj(U[j * n + k]);
        double worst = 0, lmax = 0;
        for (auto& l : lam) lmax = std::max(lmax, std::abs(l));
        if (!run(n, A, ev, V)) worst = 1e300;
        else
            for (auto& l : lam) {
                double best = 1e300;
                for (auto& e : ev) best = std::min(best, std::abs(e - l));
                worst = std::max(worst, best / lmax);
            }
        const bool pass = worst <= 1e-11;
        bad += !pass;
        std::printf("spectrum n=%d max error=%.2e %s\n", n, worst, pass ? "ok" : "FAIL");
    }
    return bad ? 1 : 0;
}
