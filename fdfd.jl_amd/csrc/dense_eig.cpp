// Small dense complex eigensolver used by the Arnoldi driver (host code).
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

namespace fdfd {
using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// small dense complex eigensolver (row-major n x n): A = Q T Q^H, then the
// eigenvectors of T by back substitution.  Householder reduction to
// Hessenberg form, implicit single-shift QR with Wilkinson shifts.
// ---------------------------------------------------------------------------
static void givens(cd x, cd y, double& c, cd& s, cd& r) {
    double ax = std::abs(x), ay = std::abs(y);
    if (ay == 0) { c = 1; s = 0; r = x; return; }
    if (ax == 0) { c = 0; s = 1; r = y; return; }
    double t = std::hypot(ax, ay);
    cd ph = x / ax;
    c = ax / t;
    s = ph * std::conj(y) / t;
    r = ph * t;
}

// returns 0 on success, -1 if the QR iteration did not converge
int small_eig(int n, std::vector<cd> A, std::vector<cd>& evals, std::vector<cd>& evecs) {
    auto at = [&](std::vector<cd>& M, int i, int j) -> cd& { return M[(size_t)i * n + j]; };
    std::vector<cd> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) at(Q, i, i) = 1.0;
    // Hessenberg
    for (int k = 0; k + 2 < n; ++k) {
        double nx = 0;
        for (int i = k + 1; i < n; ++i) nx += std::norm(at(A, i, k));
        nx = std::sqrt(nx);
        if (nx == 0) continue;
        cd x0 = at(A, k + 1, k);
        cd alpha = (std::abs(x0) > 0 ? -x0 / std::abs(x0) : cd(-1.0)) * nx;
        std::vector<cd> v(n, 0.0);
        for (int i = k + 1; i < n; ++i) v[i] = at(A, i, k);
        v[k + 1] -= alpha;
        double vn = 0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        vn = std::sqrt(vn);
        if (vn == 0) continue;
        for (int i = k + 1; i < n; ++i) v[i] /= vn;
        // A = (I - 2vv^H) A
        for (int j = 0; j < n; ++j) {
            cd s = 0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * at(A, i, j);
            for (int i = k + 1; i < n; ++i) at(A, i, j) -= 2.0 * v[i] * s;
        }
        // A = A (I - 2vv^H), Q = Q (I - 2vv^H)
        for (int i = 0; i < n; ++i) {
            cd s = 0, sq = 0;
            for (int j = k + 1; j < n; ++j) {
                s += at(A, i, j) * v[j];
                sq += at(Q, i, j) * v[j];
            }
            for (int j = k + 1; j < n; ++j) {
                at(A, i, j) -= 2.0 * s * std::conj(v[j]);
                at(Q, i, j) -= 2.0 * sq * std::conj(v[j]);
            }
        }
        for (int i = k + 2; i < n; ++i) at(A, i, k) = 0;
    }
    // QR iterations
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, iter = 0, total = 0;
    while (hi > 0) {
        int l = hi;
        while (l > 0) {
            double s = std::abs(at(A, l - 1, l - 1)) + std::abs(at(A, l, l));
            if (s == 0) s = 1;
            if (std::abs(at(A, l, l - 1)) <= eps * s) break;
            --l;
        }
        if (l > 0) at(A, l, l - 1) = 0;
        if (l == hi) { --hi; iter = 0; continue; }
        if (++total > 100 * n) return -1;
        cd a = at(A, hi - 1, hi - 1), b = at(A, hi - 1, hi), c = at(A, hi, hi - 1), d = at(A, hi, hi);
        cd mu;
        if (iter > 0 && iter % 11 == 0) {
            mu = d + std::abs(at(A, hi, hi - 1));  // exceptional shift
        } else {
            cd disc = std::sqrt((a - d) * (a - d) / 4.0 + b * c);
            cd m1 = (a + d) / 2.0 + disc, m2 = (a + d) / 2.0 - disc;
            mu = (std::abs(m1 - d) < std::abs(m2 - d)) ? m1 : m2;
        }
        ++iter;
        for (int k = l; k < hi; ++k) {
            cd x = (k == l) ? at(A, l, l) - mu : at(A, k, k - 1);
            cd y = (k == l) ? at(A, l + 1, l) : at(A, k + 1, k - 1);
            double cs; cd sn, r;
            givens(x, y, cs, sn, r);
            for (int j = (k == l ? l : k - 1); j < n; ++j) {
                cd p = at(A, k, j), q = at(A, k + 1, j);
                at(A, k, j) = cs * p + sn * q;
                at(A, k + 1, j) = -std::conj(sn) * p + cs * q;
            }
            if (k > l) at(A, k + 1, k - 1) = 0;
            const int imax = std::min(k + 2, hi);
            for (int i = 0; i <= imax; ++i) {
                cd p = at(A, i, k), q = at(A, i, k + 1);
                at(A, i, k) = p * cs + q * std::conj(sn);
                at(A, i, k + 1) = -p * sn + q * cs;
            }
            for (int i = 0; i < n; ++i) {
                cd p = at(Q, i, k), q = at(Q, i, k + 1);
                at(Q, i, k) = p * cs + q * std::conj(sn);
                at(Q, i, k + 1) = -p * sn + q * cs;
            }
        }
    }
    evals.assign(n, 0.0);
    for (int i = 0; i < n; ++i) evals[i] = at(A, i, i);
    // eigenvectors of T, then Q z
    double tnorm = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) tnorm = std::max(tnorm, std::abs(at(A, i, j)));
    const double small = eps * std::max(tnorm, 1e-300);
    evecs.assign((size_t)n * n, 0.0);
    std::vector<cd> z(n);
    for (int k = 0; k < n; ++k) {
        std::fill(z.begin(), z.end(), cd(0));
        z[k] = 1.0;
        for (int i = k - 1; i >= 0; --i) {
            cd s = 0;
            for (int j = i + 1; j <= k; ++j) s += at(A, i, j) * z[j];
            cd den = at(A, i, i) - evals[k];
            if (std::abs(den) < small) den = small;
            z[i] = -s / den;
        }
        double nz = 0;
        std::vector<cd> v(n, 0.0);
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j <= k; ++j) v[i] += at(Q, i, j) * z[j];
            nz += std::norm(v[i]);
        }
        nz = std::sqrt(nz);
        for (int i = 0; i < n; ++i) evecs[(size_t)i * n + k] = v[i] / nz;  // column k
    }
    return 0;
}

}  // namespace fdfd
