// Small dense complex eigensolver (host/device, FP64), used for the
// projected matrices of the Arnoldi drivers: the 2-D eigenfrequency driver
// on the host (eigen.cpp) and the batched 1-D mode kernel on the device
// (k_mode.hip).  Storage is supplied by the caller, so the device version
// can run on LDS-resident arrays.
//
// A = Q T Q^H: Householder reduction to Hessenberg form, then implicit
// single-shift QR with Wilkinson shifts (exceptional shift every 11th
// iteration without deflation) to upper triangular T, then the eigenvectors
// of T by back substitution, mapped back by Q and normalised to unit 2-norm.
#pragma once

#include "tri_qr.h"

namespace fdfd {
namespace eig {

using tri::cz;
using tri::mkz;
using tri::zadd;
using tri::zconj;
using tri::zmul;
using tri::zscale;
using tri::zsub;

TRI_HD inline double znorm2(cz a) { return a.re * a.re + a.im * a.im; }
TRI_HD inline double zmag(cz a) { return hypot(a.re, a.im); }
TRI_HD inline cz zdiv(cz a, cz b) { return zmul(a, tri::zinv(b)); }
// principal square root
TRI_HD inline cz zsqrt(cz a) {
    const double m = zmag(a);
    if (m == 0.0) return mkz(0, 0);
    if (a.re >= 0) {
        const double t = sqrt(0.5 * (m + a.re));
        return mkz(t, a.im / (2 * t));
    }
    const double t = sqrt(0.5 * (m - a.re));
    return mkz(fabs(a.im) / (2 * t), a.im >= 0 ? t : -t);
}

// G = [[c, s], [-conj(s), c]] with G [x; y] = [r; 0], c real
TRI_HD inline void qr_givens(cz x, cz y, double& c, cz& s, cz& r) {
    const double ax = zmag(x), ay = zmag(y);
    if (ay == 0) { c = 1; s = mkz(0, 0); r = x; return; }
    if (ax == 0) { c = 0; s = mkz(1, 0); r = y; return; }
    const double t = hypot(ax, ay);
    const cz ph = zscale(x, 1.0 / ax);
    c = ax / t;
    s = zscale(zmul(ph, zconj(y)), 1.0 / t);
    r = zscale(ph, t);
}

// Row-major n x n.  A is overwritten (ends as T).  Outputs: evals[n],
// evecs[i * n + k] = component i of eigenvector k.  Scratch: Q (n*n),
// z, v (n each).  Returns 0, or -1 if the QR iteration did not converge.
TRI_HD inline int small_eig(int n, cz* A, cz* Q, cz* evals, cz* evecs, cz* z, cz* v) {
    auto at = [n](cz* M, int i, int j) -> cz& { return M[i * n + j]; };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) at(Q, i, j) = mkz(i == j ? 1.0 : 0.0, 0);
    // Hessenberg reduction
    for (int k = 0; k + 2 < n; ++k) {
        double nx = 0;
        for (int i = k + 1; i < n; ++i) nx += znorm2(at(A, i, k));
        nx = sqrt(nx);
        if (nx == 0) continue;
        const cz x0 = at(A, k + 1, k);
        const double a0 = zmag(x0);
        const cz alpha = a0 > 0 ? zscale(x0, -nx / a0) : mkz(-nx, 0);
        for (int i = 0; i < n; ++i) v[i] = mkz(0, 0);
        for (int i = k + 1; i < n; ++i) v[i] = at(A, i, k);
        v[k + 1] = zsub(v[k + 1], alpha);
        double vn = 0;
        for (int i = k + 1; i < n; ++i) vn += znorm2(v[i]);
        vn = sqrt(vn);
        if (vn == 0) continue;
        for (int i = k + 1; i < n; ++i) v[i] = zscale(v[i], 1.0 / vn);
        // A = (I - 2 v v^H) A
        for (int j = 0; j < n; ++j) {
            cz s = mkz(0, 0);
            for (int i = k + 1; i < n; ++i) s = zadd(s, zmul(zconj(v[i]), at(A, i, j)));
            for (int i = k + 1; i < n; ++i) at(A, i, j) = zsub(at(A, i, j), zscale(zmul(v[i], s), 2.0));
        }
        // A = A (I - 2 v v^H), Q = Q (I - 2 v v^H)
        for (int i = 0; i < n; ++i) {
            cz s = mkz(0, 0), sq = mkz(0, 0);
            for (int j = k + 1; j < n; ++j) {
                s = zadd(s, zmul(at(A, i, j), v[j]));
                sq = zadd(sq, zmul(at(Q, i, j), v[j]));
            }
            for (int j = k + 1; j < n; ++j) {
                at(A, i, j) = zsub(at(A, i, j), zscale(zmul(s, zconj(v[j])), 2.0));
                at(Q, i, j) = zsub(at(Q, i, j), zscale(zmul(sq, zconj(v[j])), 2.0));
            }
        }
        for (int i = k + 2; i < n; ++i) at(A, i, k) = mkz(0, 0);
    }
    // shifted QR iterations
    const double eps = 2.220446049250313e-16;
    int hi = n - 1, iter = 0, total = 0;
    while (hi > 0) {
        int l = hi;
        while (l > 0) {
            double s = zmag(at(A, l - 1, l - 1)) + zmag(at(A, l, l));
            if (s == 0) s = 1;
            if (zmag(at(A, l, l - 1)) <= eps * s) break;
            --l;
        }
        if (l > 0) at(A, l, l - 1) = mkz(0, 0);
        if (l == hi) {
            --hi;
            iter = 0;
            continue;
        }
        if (++total > 100 * n) return -1;
        const cz a = at(A, hi - 1, hi - 1), b = at(A, hi - 1, hi), c = at(A, hi, hi - 1),
                 d = at(A, hi, hi);
        cz mu;
        if (iter > 0 && iter % 11 == 0) {
            mu = zadd(d, mkz(zmag(at(A, hi, hi - 1)), 0));  // exceptional shift
        } else {
            const cz amd = zsub(a, d);
            const cz disc = zsqrt(zadd(zscale(zmul(amd, amd), 0.25), zmul(b, c)));
            const cz half = zscale(zadd(a, d), 0.5);
            const cz m1 = zadd(half, disc), m2 = zsub(half, disc);
            mu = (zmag(zsub(m1, d)) < zmag(zsub(m2, d))) ? m1 : m2;
        }
        ++iter;
        for (int k = l; k < hi; ++k) {
            const cz x = (k == l) ? zsub(at(A, l, l), mu) : at(A, k, k - 1);
            const cz y = (k == l) ? at(A, l + 1, l) : at(A, k + 1, k - 1);
            double cs;
            cz sn, r;
            qr_givens(x, y, cs, sn, r);
            for (int j = (k == l ? l : k - 1); j < n; ++j) {
                const cz p = at(A, k, j), q = at(A, k + 1, j);
                at(A, k, j) = zadd(zscale(p, cs), zmul(sn, q));
                at(A, k + 1, j) = zsub(zscale(q, cs), zmul(zconj(sn), p));
            }
            if (k > l) at(A, k + 1, k - 1) = mkz(0, 0);
            const int imax = (k + 2 < hi) ? k + 2 : hi;
            for (int i = 0; i <= imax; ++i) {
                const cz p = at(A, i, k), q = at(A, i, k + 1);
                at(A, i, k) = zadd(zscale(p, cs), zmul(q, zconj(sn)));
                at(A, i, k + 1) = zsub(zscale(q, cs), zmul(p, sn));
            }
            for (int i = 0; i < n; ++i) {
                const cz p = at(Q, i, k), q = at(Q, i, k + 1);
                at(Q, i, k) = zadd(zscale(p, cs), zmul(q, zconj(sn)));
                at(Q, i, k + 1) = zsub(zscale(q, cs), zmul(p, sn));
            }
        }
    }
    for (int i = 0; i < n; ++i) evals[i] = at(A, i, i);
    // eigenvectors of T, then Q z
    double tnorm = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            const double m = zmag(at(A, i, j));
            tnorm = m > tnorm ? m : tnorm;
        }
    const double small = eps * (tnorm > 1e-300 ? tnorm : 1e-300);
    for (int k = 0; k < n; ++k) {
        for (int i = 0; i < n; ++i) z[i] = mkz(0, 0);
        z[k] = mkz(1, 0);
        for (int i = k - 1; i >= 0; --i) {
            cz s = mkz(0, 0);
            for (int j = i + 1; j <= k; ++j) s = zadd(s, zmul(at(A, i, j), z[j]));
            cz den = zsub(at(A, i, i), evals[k]);
            if (zmag(den) < small) den = mkz(small, 0);
            z[i] = zscale(zdiv(s, den), -1.0);
        }
        double nz = 0;
        for (int i = 0; i < n; ++i) {
            cz t = mkz(0, 0);
            for (int j = 0; j <= k; ++j) t = zadd(t, zmul(at(Q, i, j), z[j]));
            v[i] = t;
            nz += znorm2(t);
        }
        nz = sqrt(nz);
        for (int i = 0; i < n; ++i) evecs[i * n + k] = zscale(v[i], 1.0 / nz);
    }
    return 0;
}

}  // namespace eig
}  // namespace fdfd
