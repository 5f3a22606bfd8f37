// Inverse of a periodic tridiagonal line block by Givens QR (host/device,
// FP64 arithmetic).  Used for the lines eliminated on the first reduction
// level, whose block is the line's own operator
//   T[i][i] = d[i], T[i][i+1 mod n] = e[i], T[i][i-1 mod n] = f[i].
//
// Rotations, in order, for k = 0 .. n-2:
//   rot1 on rows (k, k+1)  (rows (n-2, n-1) for k = n-2)  zeroes T[k+1][k]
//   rot2 on rows (k, n-1)  (k <= n-3 only)                zeroes T[n-1][k]
// leave R = Q T upper triangular with two superdiagonals plus the last two
// columns (the wrap corner fills them).  No pivoting decision is taken and
// the factorization is backward stable.  X = T^-1 = R^-1 Q is then formed
// one column at a time: column b is Q e_b (the rotations streamed over the
// column, O(n)) followed by back substitution with R (O(n)).
//
// Work array (doubles) of a line of n cells:
//   rotations, k = 0..n-1 (6 each): c1, s1.re, s1.im, c2, s2.re, s2.im
//   rows of R, k = 0..n-1 (10 each): 1/R[k][k], R[k][k+1], R[k][k+2],
//     R[k][n-2], R[k][n-1] (complex).  Entries that fall in the last two
//     columns are kept only in the last two slots.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define TRI_HD __host__ __device__
#else
#define TRI_HD
#endif

namespace fdfd {
namespace tri {

struct cz {
    double re, im;
};
TRI_HD inline cz mkz(double r, double i) {
    cz z;
    z.re = r;
    z.im = i;
    return z;
}
TRI_HD inline cz zadd(cz a, cz b) { return mkz(a.re + b.re, a.im + b.im); }
TRI_HD inline cz zsub(cz a, cz b) { return mkz(a.re - b.re, a.im - b.im); }
TRI_HD inline cz zmul(cz a, cz b) { return mkz(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
TRI_HD inline cz zscale(cz a, double s) { return mkz(a.re * s, a.im * s); }
TRI_HD inline cz zconj(cz a) { return mkz(a.re, -a.im); }
TRI_HD inline double zabs(cz a) { return hypot(a.re, a.im); }
TRI_HD inline cz zinv(cz z) {  // Smith's method
    if (fabs(z.re) >= fabs(z.im)) {
        const double t = z.im / z.re, d = z.re + z.im * t;
        return mkz(1.0 / d, -t / d);
    }
    const double t = z.re / z.im, d = z.im + z.re * t;
    return mkz(t / d, -1.0 / d);
}

// G = [[c, s], [-conj(s), c]] (c real, unitary) with G [a; b] = [r; 0]
TRI_HD inline void givens(cz a, cz b, double& c, cz& s, cz& r) {
    const double aa = zabs(a), ab = zabs(b);
    if (ab == 0.0) {
        c = 1.0;
        s = mkz(0, 0);
        r = a;
        return;
    }
    if (aa == 0.0) {
        c = 0.0;
        s = zscale(zconj(b), 1.0 / ab);
        r = mkz(ab, 0);
        return;
    }
    const double t = hypot(aa, ab);
    const cz ph = zscale(a, 1.0 / aa);
    c = aa / t;
    s = zscale(zmul(ph, zconj(b)), 1.0 / t);
    r = zscale(ph, t);
}
// (x, y) <- G (x, y)
TRI_HD inline void rot(double c, cz s, cz& x, cz& y) {
    const cz nx = zadd(zscale(x, c), zmul(s, y));
    const cz ny = zsub(zscale(y, c), zmul(zconj(s), x));
    x = nx;
    y = ny;
}

constexpr int ROT = 6, RROW = 10;
TRI_HD inline int work_doubles(int n) { return (ROT + RROW) * n; }
TRI_HD inline void put_rot(double* w, int k, int which, double c, cz s) {
    double* p = w + ROT * k + 3 * which;
    p[0] = c;
    p[1] = s.re;
    p[2] = s.im;
}
TRI_HD inline void get_rot(const double* w, int k, int which, double& c, cz& s) {
    const double* p = w + ROT * k + 3 * which;
    c = p[0];
    s = mkz(p[1], p[2]);
}
// slot 0: 1/R[k][k], 1: R[k][k+1], 2: R[k][k+2], 3: R[k][n-2], 4: R[k][n-1]
TRI_HD inline void put_r(double* w, int n, int k, int slot, cz v) {
    double* p = w + ROT * n + RROW * k + 2 * slot;
    p[0] = v.re;
    p[1] = v.im;
}
TRI_HD inline cz get_r(const double* w, int n, int k, int slot) {
    const double* p = w + ROT * n + RROW * k + 2 * slot;
    return mkz(p[0], p[1]);
}

// Dense Givens QR of the trailing s x s part (rows and columns g0 .. n-1,
// g0 + s == n), with the same rotation order as the banded part.
TRI_HD inline void dense_tail(int n, int g0, int s, cz (*M)[4], double* w, double& pmin,
                              double& pmax) {
    for (int kk = 0; kk + 1 < s; ++kk) {
        const int np = (kk + 1 < s - 1) ? 2 : 1;
        for (int q = 0; q < np; ++q) {
            const int p = (q == 0 && kk + 1 < s - 1) ? kk + 1 : s - 1;
            double c;
            cz sn, r;
            givens(M[kk][kk], M[p][kk], c, sn, r);
            M[kk][kk] = r;
            M[p][kk] = mkz(0, 0);
            for (int j = kk + 1; j < s; ++j) rot(c, sn, M[kk][j], M[p][j]);
            put_rot(w, g0 + kk, q, c, sn);
        }
        if (np == 1) put_rot(w, g0 + kk, 1, 1.0, mkz(0, 0));
    }
    for (int i = 0; i < s; ++i) {
        const int k = g0 + i;
        const double a = zabs(M[i][i]);
        pmin = a < pmin ? a : pmin;
        pmax = a > pmax ? a : pmax;
        put_r(w, n, k, 0, a > 0 ? zinv(M[i][i]) : mkz(0, 0));
        for (int slot = 1; slot < 5; ++slot) put_r(w, n, k, slot, mkz(0, 0));
        for (int j = i + 1; j < s; ++j) {
            const int col = g0 + j;
            const int slot = (col == n - 1) ? 4 : (col == n - 2) ? 3 : (col == k + 1) ? 1 : 2;
            put_r(w, n, k, slot, M[i][j]);
        }
    }
}

// Factor the block given by get(i, d, e, f).  Returns 0, or -1 if R has a
// zero diagonal entry (singular block).  pmin / pmax: smallest and largest
// |R[k][k]|.
template <typename Get>
TRI_HD inline int factor(int n, Get get, double* w, double& pmin, double& pmax) {
    pmin = 1e300;
    pmax = 0.0;
    cz M[4][4];
    if (n <= 4) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) M[i][j] = mkz(0, 0);
        for (int i = 0; i < n; ++i) {
            cz d, e, f;
            get(i, d, e, f);
            M[i][i] = zadd(M[i][i], d);
            M[i][(i + 1) % n] = zadd(M[i][(i + 1) % n], e);
            M[i][(i + n - 1) % n] = zadd(M[i][(i + n - 1) % n], f);
        }
        dense_tail(n, 0, n, M, w, pmin, pmax);
        return pmin > 0 ? 0 : -1;
    }
    cz d, e, f;
    get(0, d, e, f);
    // current row k at columns k, k+1, k+2, n-2, n-1
    cz c0 = d, c1 = e, c2 = mkz(0, 0), cA = mkz(0, 0), cB = f;
    get(n - 1, d, e, f);
    // last row at columns k, k+1, n-2, n-1
    cz L0 = e, L1 = mkz(0, 0), LA = f, LB = d;
    for (int k = 0; k + 5 <= n; ++k) {
        cz dk, ek, fk;
        get(k + 1, dk, ek, fk);  // row k+1: fk at k, dk at k+1, ek at k+2
        double c;
        cz s, r;
        givens(c0, fk, c, s, r);
        put_rot(w, k, 0, c, s);
        cz n1 = c1, m1 = dk;
        rot(c, s, n1, m1);
        cz n2 = c2, m2 = ek;
        rot(c, s, n2, m2);
        cz nA = cA, mA = mkz(0, 0);
        rot(c, s, nA, mA);
        cz nB = cB, mB = mkz(0, 0);
        rot(c, s, nB, mB);
        // rows (k, n-1): zero L0
        double c2r;
        cz s2, r2;
        givens(r, L0, c2r, s2, r2);
        put_rot(w, k, 1, c2r, s2);
        cz l1 = L1;
        rot(c2r, s2, n1, l1);
        cz z2 = mkz(0, 0);
        rot(c2r, s2, n2, z2);
        rot(c2r, s2, nA, LA);
        rot(c2r, s2, nB, LB);
        const double a = zabs(r2);
        pmin = a < pmin ? a : pmin;
        pmax = a > pmax ? a : pmax;
        put_r(w, n, k, 0, a > 0 ? zinv(r2) : mkz(0, 0));
        put_r(w, n, k, 1, n1);
        put_r(w, n, k, 2, n2);
        put_r(w, n, k, 3, nA);
        put_r(w, n, k, 4, nB);
        // next current row = row k+1 after rot1; next last row
        c0 = m1;
        c1 = m2;
        c2 = mkz(0, 0);
        cA = mA;
        cB = mB;
        L0 = l1;
        L1 = z2;
    }
    // rows n-4 .. n-1, columns n-4 .. n-1
    const int g0 = n - 4;
    M[0][0] = c0; M[0][1] = c1; M[0][2] = cA; M[0][3] = cB;
    get(n - 3, d, e, f);
    M[1][0] = f; M[1][1] = d; M[1][2] = e; M[1][3] = mkz(0, 0);
    get(n - 2, d, e, f);
    M[2][0] = mkz(0, 0); M[2][1] = f; M[2][2] = d; M[2][3] = e;
    M[3][0] = L0; M[3][1] = L1; M[3][2] = LA; M[3][3] = LB;
    dense_tail(n, g0, 4, M, w, pmin, pmax);
    return pmin > 0 ? 0 : -1;
}

// Column b of X = T^-1.  First pass: y = Q e_b through put(k, y[k]) for
// k >= first (y[k] = 0 below); second pass: x = R^-1 y, reading y[k] with
// get(k) for k >= first and writing x[k] with put(k, x[k]) for every k.
TRI_HD inline int first_nonzero(int n, int b) { return (b == n - 1) ? 0 : (b > 0 ? b - 1 : 0); }

template <typename Put>
TRI_HD inline void apply_q(int n, const double* w, int b, int kstart, Put put) {
    // state: carry = v[k], tail = v[n-1]
    const int k0 = first_nonzero(n, b);
    cz carry = mkz(b == 0 ? 1.0 : 0.0, 0), tail = mkz(b == n - 1 ? 1.0 : 0.0, 0);
    if (n == 1) {
        put(0, mkz(1, 0));
        return;
    }
    for (int k = kstart; k + 2 < n; ++k) {  // k <= n-3: rot1 (k, k+1), rot2 (k, n-1)
        if (k < k0) continue;
        cz nxt = mkz((k + 1 == b) ? 1.0 : 0.0, 0);
        double c;
        cz s;
        get_rot(w, k, 0, c, s);
        rot(c, s, carry, nxt);
        get_rot(w, k, 1, c, s);
        rot(c, s, carry, tail);
        put(k, carry);
        carry = nxt;
    }
    double c;
    cz s;
    get_rot(w, n - 2, 0, c, s);  // rot1 on rows (n-2, n-1)
    rot(c, s, carry, tail);
    put(n - 2, carry);
    put(n - 1, tail);
}

// y = Q v for a general vector v (get(k) = v[k]); y written through
// put(k, y[k]) in increasing k.  put may overwrite v[k] in place: v[k] is
// read before y[k] is written, and never after.
template <typename Get, typename Put>
TRI_HD inline void apply_q_vec(int n, const double* w, Get get, Put put) {
    if (n == 1) {
        put(0, get(0));
        return;
    }
    cz carry = get(0), tail = get(n - 1);
    for (int k = 0; k + 2 < n; ++k) {
        cz nxt = get(k + 1);
        double c;
        cz s;
        get_rot(w, k, 0, c, s);
        rot(c, s, carry, nxt);
        get_rot(w, k, 1, c, s);
        rot(c, s, carry, tail);
        put(k, carry);
        carry = nxt;
    }
    double c;
    cz s;
    get_rot(w, n - 2, 0, c, s);
    rot(c, s, carry, tail);
    put(n - 2, carry);
    put(n - 1, tail);
}

template <typename Get, typename Put>
TRI_HD inline void apply_rinv(int n, const double* w, int b, Get get, Put put) {
    const int k0 = first_nonzero(n, b);
    auto y = [&](int k) { return k >= k0 ? get(k) : mkz(0, 0); };
    const cz xB = zmul(y(n - 1), get_r(w, n, n - 1, 0));
    put(n - 1, xB);
    if (n == 1) return;
    const cz xA = zmul(zsub(y(n - 2), zmul(get_r(w, n, n - 2, 4), xB)), get_r(w, n, n - 2, 0));
    put(n - 2, xA);
    cz x1 = xA, x2 = xB;  // x[k+1], x[k+2]
    for (int k = n - 3; k >= 0; --k) {
        cz v = y(k);
        v = zsub(v, zmul(get_r(w, n, k, 1), x1));
        v = zsub(v, zmul(get_r(w, n, k, 2), x2));
        v = zsub(v, zmul(get_r(w, n, k, 3), xA));
        v = zsub(v, zmul(get_r(w, n, k, 4), xB));
        const cz x = zmul(v, get_r(w, n, k, 0));
        put(k, x);
        x2 = x1;
        x1 = x;
    }
}

}  // namespace tri
}  // namespace fdfd
