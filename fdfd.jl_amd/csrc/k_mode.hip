// Batched 1-D eigenmodes for mode sources: eigenmode(d::Device{1}, pol,
// neff, nev) of src/solver/eigen.jl:6-29, called by get_modes / setup_mode!
// (src/device.jl:118-149) on a slice of eps.
//
// Operator of a slice of n cells (periodic, no PML; grid.jl:138,141):
//   TM: A = k0^2 diag(eps) + dxf dxb
//       A[i][i +- 1] = 1/dx^2, A[i][i] = k0^2 eps_i - 2/dx^2
//   TE: A = k0^2 diag(eps) + diag(eps) dxf diag(1/avg eps) dxb
//       A[i][i+1] = eps_i / (dx^2 a_{i+1}), A[i][i-1] = eps_i / (dx^2 a_i),
//       A[i][i] = k0^2 eps_i - A[i][i+1] - A[i][i-1],  a_i = (eps_i + eps_{i-1})/2
// (grid_average, grid.jl:157-158), k0^2 = omega^2 mu0' eps0'.
//
// Eigs with sigma (eigen.jl:26) is shift-invert Arnoldi on
// OP = (A - sigma I)^-1 keeping the nev Ritz values of largest magnitude,
// i.e. the eigenvalues nearest sigma.  One workgroup per problem:
//   * A - sigma I is factored once by Givens QR (tri_qr.h, one lane; the
//     operator is periodic tridiagonal, O(n)), and every OP application is
//     one O(n) solve on one lane;
//   * Arnoldi vectors live in global memory; CGS2 dot products, updates and
//     norms use the whole workgroup;
//   * the projected (ncv x ncv) eigenproblem runs on one lane from LDS
//     (small_eig.h, the solver of the 2-D eigenfrequency driver);
//   * thick restart keeps the span of the `keep` wanted Ritz vectors.
// Output: beta = sqrt(sigma + 1/theta) (principal branch, eigen.jl:28) nearest
// first, eigenvectors of unit 2-norm with their largest entry real > 0.
#include "fdfd_common.h"
#include "small_eig.h"
#include "tri_qr.h"

namespace fdfd {

namespace {

constexpr int MT = 256;       // threads per workgroup
constexpr int MW = MT / 64;   // waves per workgroup

using eig::cz;
using eig::mkz;

__device__ inline cz ld(const double2* p) {
    const double2 v = *p;
    return mkz(v.x, v.y);
}
__device__ inline void st(double2* p, cz v) { *p = make_double2(v.re, v.im); }

__device__ inline double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

__global__ __launch_bounds__(MT) void k_mode(const ModeJob* jobs, int nev, double tol,
                                             int maxiter) {
    const ModeJob J = jobs[blockIdx.x];
    const int n = J.n, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int m = J.ncv;
    __shared__ cz H[(MODE_NCV + 1) * MODE_NCV];  // (m + 1) x m, row-major
    __shared__ cz S0[MODE_NCV * MODE_NCV];       // small_eig input, then Qk (m x keep)
    __shared__ cz S1[MODE_NCV * MODE_NCV];       // small_eig Q, then H Qk
    __shared__ cz Y[MODE_NCV * MODE_NCV];        // Ritz vectors of the projected matrix
    __shared__ cz theta[MODE_NCV], zs[MODE_NCV], vs[MODE_NCV + 1];
    __shared__ cz hc[MODE_NCV + 1];
    __shared__ double rre[MODE_NCV + 1][MW], rim[MODE_NCV + 1][MW];
    __shared__ double rmag[MW];
    __shared__ int ridx[MW];
    __shared__ int order[MODE_NCV];
    __shared__ int flag[4];  // status, nconv, done, restarts
    const int keep = min(m - 1, nev + max(2, (m - nev) / 2));
    auto vp = [&](int q) { return J.V + (size_t)q * n; };

    // out[q] = V_q^H w for q < cnt (all threads; out valid after return)
    auto dots = [&](int cnt, const double2* w, cz* out) {
        for (int q = 0; q < cnt; ++q) {
            const double2* v = vp(q);
            double sr = 0, si = 0;
            for (int i = t; i < n; i += MT) {
                const double2 a = v[i], b = w[i];
                sr += a.x * b.x + a.y * b.y;
                si += a.x * b.y - a.y * b.x;
            }
            sr = wsum(sr);
            si = wsum(si);
            if (lane == 0) {
                rre[q][wv] = sr;
                rim[q][wv] = si;
            }
        }
        __syncthreads();
        if (t < cnt) {
            double sr = 0, si = 0;
            for (int k = 0; k < MW; ++k) {
                sr += rre[t][k];
                si += rim[t][k];
            }
            out[t] = mkz(sr, si);
        }
        __syncthreads();
    };
    // w -= sum_q V_q c[q]
    auto update = [&](int cnt, double2* w, const cz* c) {
        for (int i = t; i < n; i += MT) {
            double xr = w[i].x, xi = w[i].y;
            for (int q = 0; q < cnt; ++q) {
                const double2 a = vp(q)[i];
                xr -= a.x * c[q].re - a.y * c[q].im;
                xi -= a.x * c[q].im + a.y * c[q].re;
            }
            w[i] = make_double2(xr, xi);
        }
        __syncthreads();
    };
    auto norm = [&](const double2* w) -> double {
        double s = 0;
        for (int i = t; i < n; i += MT) s += w[i].x * w[i].x + w[i].y * w[i].y;
        s = wsum(s);
        if (lane == 0) rre[0][wv] = s;
        __syncthreads();
        double tot = 0;
        for (int k = 0; k < MW; ++k) tot += rre[0][k];
        __syncthreads();
        return sqrt(tot);
    };
    auto scale = [&](double2* w, cz a) {
        for (int i = t; i < n; i += MT) {
            const double2 x = w[i];
            w[i] = make_double2(x.x * a.re - x.y * a.im, x.x * a.im + x.y * a.re);
        }
        __syncthreads();
    };
    int applies = 0;
    // w = OP v: copy, then one lane solves in place with the QR factors
    auto op = [&](const double2* v, double2* w) {
        for (int i = t; i < n; i += MT) w[i] = v[i];
        __syncthreads();
        if (t == 0) {
            auto get = [&](int k) { return ld(w + k); };
            auto put = [&](int k, cz y) { st(w + k, y); };
            tri::apply_q_vec(n, J.trw, get, put);
            tri::apply_rinv(n, J.trw, 0, get, put);
        }
        ++applies;
        __syncthreads();
    };

    // ---- factor A - sigma I (one lane) and the start vector --------------
    if (t == 0) {
        const cz sg = mkz(J.sigma.x, J.sigma.y);
        auto coef = [&](int i, cz& d, cz& e, cz& f) {
            const cz ei = ld(J.eps + i);
            const cz k2e = eig::zscale(ei, J.k0sq);
            if (J.pol == 1) {
                e = mkz(J.idx2, 0);
                f = mkz(J.idx2, 0);
            } else {
                const cz em = ld(J.eps + (i + n - 1) % n), ep = ld(J.eps + (i + 1) % n);
                const cz a_i = eig::zscale(eig::zadd(ei, em), 0.5);
                const cz a_ip = eig::zscale(eig::zadd(ep, ei), 0.5);
                e = eig::zscale(eig::zdiv(ei, a_ip), J.idx2);
                f = eig::zscale(eig::zdiv(ei, a_i), J.idx2);
            }
            d = eig::zsub(eig::zsub(eig::zsub(k2e, e), f), sg);
        };
        double pmin, pmax;
        const int rc = tri::factor(n, coef, J.trw, pmin, pmax);
        flag[0] = rc ? MODE_SINGULAR : MODE_OK;
        flag[1] = 0;
        flag[2] = 0;
        flag[3] = 0;
        // deterministic start vector (xorshift64)
        unsigned long long s = 0x9e3779b97f4a7c15ULL;
        double2* v0 = vp(0);
        for (int i = 0; i < n; ++i) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const double a = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const double b = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            v0[i] = make_double2(a, b);
        }
    }
    __syncthreads();
    if (flag[0] != MODE_OK) {
        if (t == 0) {
            J.info[0] = flag[0];
            J.info[1] = 0;
            J.info[2] = 0;
            J.info[3] = 0;
        }
        return;  // uniform: every thread read the same flag
    }
    {
        const double nv = norm(vp(0));
        scale(vp(0), mkz(1.0 / nv, 0));
    }
    for (int i = t; i < (m + 1) * m; i += MT) H[i] = mkz(0, 0);
    __syncthreads();

    // ---- thick-restart Arnoldi -------------------------------------------
    int start = 0;
    for (;;) {
        for (int j = start; j < m; ++j) {
            double2* w = vp(j + 1);
            op(vp(j), w);
            dots(j + 1, w, hc);
            if (t <= j) H[t * m + j] = hc[t];
            __syncthreads();
            update(j + 1, w, hc);
            dots(j + 1, w, hc);
            if (t <= j) H[t * m + j] = eig::zadd(H[t * m + j], hc[t]);
            __syncthreads();
            update(j + 1, w, hc);
            const double hn = norm(w);
            if (t == 0) H[(j + 1) * m + j] = mkz(hn, 0);
            if (hn > 0) scale(w, mkz(1.0 / hn, 0));
            __syncthreads();
        }
        if (t == 0) {
            for (int i = 0; i < m; ++i)
                for (int k = 0; k < m; ++k) S0[i * m + k] = H[i * m + k];
            const int rc = eig::small_eig(m, S0, S1, theta, Y, zs, vs);
            if (rc) {
                flag[0] = MODE_EIG_FAIL;
                flag[2] = 1;
            } else {
                // order by |theta| descending (stable insertion sort)
                for (int i = 0; i < m; ++i) order[i] = i;
                for (int i = 1; i < m; ++i) {
                    const int v = order[i];
                    const double av = eig::zmag(theta[v]);
                    int k = i - 1;
                    while (k >= 0 && eig::zmag(theta[order[k]]) < av) {
                        order[k + 1] = order[k];
                        --k;
                    }
                    order[k + 1] = v;
                }
                // residual of Ritz pair i: |H[m, :] y_i|
                int nc = 0;
                for (int q = 0; q < nev; ++q) {
                    const int i = order[q];
                    cz r = mkz(0, 0);
                    for (int k = 0; k < m; ++k) r = eig::zadd(r, eig::zmul(H[m * m + k], Y[k * m + i]));
                    if (eig::zmag(r) <= tol * eig::zmag(theta[i])) ++nc;
                }
                flag[1] = nc;
                flag[2] = (nc >= nev || flag[3] >= maxiter) ? 1 : 0;
            }
        }
        __syncthreads();
        if (flag[2]) break;
        // restart: Qk = orthonormal basis of the keep wanted Ritz vectors
        if (t == 0) {
            flag[3] += 1;
            for (int q = 0; q < keep; ++q)
                for (int i = 0; i < m; ++i) S0[i * keep + q] = Y[i * m + order[q]];
            for (int q = 0; q < keep; ++q) {
                for (int pass = 0; pass < 2; ++pass)
                    for (int p = 0; p < q; ++p) {
                        cz s = mkz(0, 0);
                        for (int i = 0; i < m; ++i)
                            s = eig::zadd(s, eig::zmul(eig::zconj(S0[i * keep + p]), S0[i * keep + q]));
                        for (int i = 0; i < m; ++i)
                            S0[i * keep + q] = eig::zsub(S0[i * keep + q], eig::zmul(s, S0[i * keep + p]));
                    }
                double nq = 0;
                for (int i = 0; i < m; ++i) nq += eig::znorm2(S0[i * keep + q]);
                nq = sqrt(nq);
                for (int i = 0; i < m; ++i) S0[i * keep + q] = eig::zscale(S0[i * keep + q], 1.0 / nq);
            }
        }
        __syncthreads();
        // W[:, q] = V_m Qk[:, q]
        for (int i = t; i < n; i += MT)
            for (int q = 0; q < keep; ++q) {
                double xr = 0, xi = 0;
                for (int k = 0; k < m; ++k) {
                    const double2 a = vp(k)[i];
                    const cz c = S0[k * keep + q];
                    xr += a.x * c.re - a.y * c.im;
                    xi += a.x * c.im + a.y * c.re;
                }
                J.W[(size_t)q * n + i] = make_double2(xr, xi);
            }
        __syncthreads();
        // v_keep = v_m, then V[:, 0:keep] = W
        for (int i = t; i < n; i += MT) vp(keep)[i] = vp(m)[i];
        __syncthreads();
        for (int i = t; i < n; i += MT)
            for (int q = 0; q < keep; ++q) vp(q)[i] = J.W[(size_t)q * n + i];
        // H_new = Qk^H H_m Qk, row keep = H[m, :] Qk
        if (t == 0) {
            for (int i = 0; i < m; ++i)
                for (int q = 0; q < keep; ++q) {
                    cz s = mkz(0, 0);
                    for (int k = 0; k < m; ++k) s = eig::zadd(s, eig::zmul(H[i * m + k], S0[k * keep + q]));
                    S1[i * keep + q] = s;
                }
            for (int q = 0; q < keep; ++q) {
                cz s = mkz(0, 0);
                for (int k = 0; k < m; ++k) s = eig::zadd(s, eig::zmul(H[m * m + k], S0[k * keep + q]));
                vs[q] = s;
            }
            for (int i = 0; i < (m + 1) * m; ++i) H[i] = mkz(0, 0);
            for (int p = 0; p < keep; ++p)
                for (int q = 0; q < keep; ++q) {
                    cz s = mkz(0, 0);
                    for (int i = 0; i < m; ++i)
                        s = eig::zadd(s, eig::zmul(eig::zconj(S0[i * keep + p]), S1[i * keep + q]));
                    H[p * m + q] = s;
                }
            for (int q = 0; q < keep; ++q) H[keep * m + q] = vs[q];
        }
        __syncthreads();
        start = keep;
    }

    // ---- outputs ------------------------------------------------------------
    const bool ok = flag[0] == MODE_OK;
    if (ok) {
        for (int q = 0; q < nev; ++q) {
            const int sel = order[q];
            double2* x = J.vec + (size_t)q * n;
            for (int i = t; i < n; i += MT) {
                double xr = 0, xi = 0;
                for (int k = 0; k < m; ++k) {
                    const double2 a = vp(k)[i];
                    const cz c = Y[k * m + sel];
                    xr += a.x * c.re - a.y * c.im;
                    xi += a.x * c.im + a.y * c.re;
                }
                x[i] = make_double2(xr, xi);
            }
            __syncthreads();
            // largest entry (lowest index among equals) and the 2-norm
            double best = -1, s2 = 0;
            int bi = 0;
            for (int i = t; i < n; i += MT) {
                const double a2 = x[i].x * x[i].x + x[i].y * x[i].y;
                s2 += a2;
                if (a2 > best) {
                    best = a2;
                    bi = i;
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double ob = __shfl_xor(best, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            s2 = wsum(s2);
            if (lane == 0) {
                rmag[wv] = best;
                ridx[wv] = bi;
                rre[0][wv] = s2;
            }
            __syncthreads();
            double gb = rmag[0], tot = 0;
            int gi = ridx[0];
            for (int k = 0; k < MW; ++k) {
                tot += rre[0][k];
                if (rmag[k] > gb || (rmag[k] == gb && ridx[k] < gi)) {
                    gb = rmag[k];
                    gi = ridx[k];
                }
            }
            const cz ph = ld(x + gi);
            const double pm = eig::zmag(ph);
            __syncthreads();
            const cz a = eig::zscale(eig::zconj(ph), 1.0 / (pm * sqrt(tot)));
            scale(x, a);
            if (t == 0) {
                const cz lam = eig::zadd(mkz(J.sigma.x, J.sigma.y), tri::zinv(theta[sel]));
                st(J.beta + q, eig::zsqrt(lam));
            }
        }
    }
    if (t == 0) {
        J.info[0] = flag[0];
        J.info[1] = ok ? flag[1] : 0;
        J.info[2] = flag[3];
        J.info[3] = applies;
    }
}

void launch_mode(hipStream_t s, const ModeJob* jobs_d, int njobs, int nev, double tol,
                 int maxiter) {
    k_mode<<<njobs, MT, 0, s>>>(jobs_d, nev, tol, maxiter);
}

}  // namespace fdfd
