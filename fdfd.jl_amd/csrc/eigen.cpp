// Shift-invert Arnoldi for fdfd_eigenfrequency (src/solver/eigen.jl:69-115).
//
// The reference calls Arpack.eigs(A, nev, sigma, which) (eigen.jl:86,104),
// which factors A - sigma*I once and runs implicitly restarted Arnoldi on
// OP = (A - sigma I)^-1.  Here:
//   TM: A = diag(1/eps) L,  L = dxf dxb + dyf dyb (stretched),
//       sigma = -w0^2 mu0' eps0'.  A - sigma I = diag(1/eps) mu0' A_TM(w0),
//       so OP x = dx^2 A'^-1 (eps .* x) with A' the scaled driven operator.
//   TE: A - sigma I = A_TE(w0), OP x = eps0' dx^2 A'^-1 x.
// A' is factored once (block factors on the GPU) and every OP application is
// an FGMRES-refined solve on those factors.  Arnoldi vectors live on the
// device (FP64, CGS2 orthogonalisation); the small projected matrix is
// handled on the host with a dense complex eigensolver.  Restarts keep the
// span of the wanted Ritz vectors (thick restart).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "../../include/fdfd.h"
#include "runtime.h"

namespace fdfd {
int rt_gmres(Handle* h, Factor* F, int trans, const double2* b, double2* x, double tol, int* iters,
             double* resid);
int rt_build_factor(Handle* h, Factor* F);
int rt_norm(Handle* h, const double2* x, long N, double* out);
int rt_multidot(Handle* h, const double2* V, long N, int k, const double2* w, cd* out);
int rt_update(Handle* h, double2* w, const double2* V, long N, int k, const cd* c, double sign);
int small_eig(int n, std::vector<cd> A, std::vector<cd>& evals, std::vector<cd>& evecs);

}  // namespace fdfd

using namespace fdfd;

extern "C" int fdfd_eigenfrequency(fdfd_handle* hh, int pol, double omega0, int nev, int ncv,
                                   int which, int precision, double tol, int maxiter,
                                   fdfd_complex* omegas, fdfd_complex* fields,
                                   fdfd_stats* stats) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !omegas || nev < 1) return FDFD_ERR_INVALID;
    if (!h->has_grid || !h->has_eps)
        return fail(h, FDFD_ERR_STATE, "grid and eps must be set before fdfd_eigenfrequency");
    if (pol != FDFD_TM && pol != FDFD_TE) return fail(h, FDFD_ERR_INVALID, "pol must be TM or TE");
    if (!(omega0 > 0)) return fail(h, FDFD_ERR_INVALID, "omega0 must be > 0");
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    if (ncv <= 0) ncv = std::max(20, 2 * nev + 1);  // Arpack.jl default
    ncv = std::min<long>(ncv, N);
    if (nev >= ncv) return fail(h, FDFD_ERR_INVALID, "nev must be < ncv (and < Nx*Ny)");
    if (ncv > 64) return fail(h, FDFD_ERR_INVALID, "ncv must be <= 64");
    if (!(tol > 0)) tol = 1e-12;
    if (maxiter <= 0) maxiter = 300;
    (void)hipSetDevice(h->device);
    const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;

    Factor* F = new Factor();
    F->pol = pol;
    F->omega = omega0;
    F->precision = precision;
    int rc = rt_build_factor(h, F);
    if (rc) { rt_free_factor(h, F); return rc; }
    const double sigma = (pol == FDFD_TM) ? -omega0 * omega0 * mu0n * eps0n : -omega0 * omega0 * mu0n;
    const double opscale = (pol == FDFD_TM) ? g.dx * g.dx : eps0n * g.dx * g.dx;
    // accuracy of each shift-invert application (FP64 FGMRES on the factors)
    const double inner_tol = std::max(1e-13, std::min(1e-12, tol * 1e-1));

    double2* Vb = nullptr;  // ncv + 1 Arnoldi vectors
    double2* Wb = nullptr;  // restart workspace
    auto cleanup = [&]() {
        if (Vb) (void)hipFree(Vb);
        if (Wb) (void)hipFree(Wb);
        rt_free_factor(h, F);
    };
    if (hipMalloc(&Vb, sizeof(double2) * N * (ncv + 1)) != hipSuccess ||
        hipMalloc(&Wb, sizeof(double2) * N * ncv) != hipSuccess) {
        cleanup();
        return fail(h, FDFD_ERR_OOM, "Arnoldi basis allocation failed");
    }
    int iters = 0;
    auto vec = [&](int i) { return Vb + (long)i * N; };
    // OP y = x
    auto op = [&](const double2* x, double2* y) -> int {
        if (pol == FDFD_TM) launch_mul_eps(h->stream, h->tmp, x, h->eps, make_double2(opscale, 0), N);
        else launch_scale_copy(h->stream, h->tmp, x, make_double2(opscale, 0), N);
        launch_zero(h->stream, y, N);
        double res = 0;
        return rt_gmres(h, F, 0, h->tmp, y, inner_tol, &iters, &res);
    };
    // deterministic start vector
    {
        std::vector<double2> v0(N);
        uint64_t s = 0x9e3779b97f4a7c15ULL;
        for (long i = 0; i < N; ++i) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            double a = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            double b = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            v0[i] = make_double2(a, b);
        }
        if ((rc = stream_copy(h, vec(0), v0.data(), sizeof(double2) * N, hipMemcpyHostToDevice))) {
            cleanup();
            return rc;
        }
        double nv;
        if ((rc = rt_norm(h, vec(0), N, &nv))) { cleanup(); return rc; }
        launch_scale_copy(h->stream, vec(0), vec(0), make_double2(1.0 / nv, 0), N);
    }
    const int m = ncv;
    std::vector<cd> H((size_t)(m + 1) * m, 0.0);  // row-major (m+1) x m
    auto Hat = [&](int i, int j) -> cd& { return H[(size_t)i * m + j]; };
    std::vector<cd> hc(m + 1), hc2(m + 1);
    int start = 0, nconv = 0, restarts = 0;
    std::vector<cd> theta, Y;
    std::vector<int> order;
    const int keep = std::min(m - 1, nev + std::max(2, (m - nev) / 2));
    for (;;) {
        for (int j = start; j < m; ++j) {
            double2* w = vec(j + 1);
            if ((rc = op(vec(j), w))) { cleanup(); return rc; }
            if ((rc = rt_multidot(h, Vb, N, j + 1, w, hc.data()))) { cleanup(); return rc; }
            if ((rc = rt_update(h, w, Vb, N, j + 1, hc.data(), -1.0))) { cleanup(); return rc; }
            if ((rc = rt_multidot(h, Vb, N, j + 1, w, hc2.data()))) { cleanup(); return rc; }
            if ((rc = rt_update(h, w, Vb, N, j + 1, hc2.data(), -1.0))) { cleanup(); return rc; }
            for (int i = 0; i <= j; ++i) Hat(i, j) = hc[i] + hc2[i];
            double hn;
            if ((rc = rt_norm(h, w, N, &hn))) { cleanup(); return rc; }
            Hat(j + 1, j) = hn;
            if (hn > 0) launch_scale_copy(h->stream, w, w, make_double2(1.0 / hn, 0), N);
        }
        std::vector<cd> Hm((size_t)m * m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) Hm[(size_t)i * m + j] = Hat(i, j);
        if (small_eig(m, Hm, theta, Y) != 0) {
            cleanup();
            return fail(h, FDFD_ERR_NOCONV, "projected eigenproblem did not converge");
        }
        order.resize(m);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (which == FDFD_WHICH_LR) return theta[a].real() > theta[b].real();
            return std::abs(theta[a]) > std::abs(theta[b]);
        });
        // residual of Ritz pair i: |H[m, :] . y_i|
        nconv = 0;
        for (int q = 0; q < nev; ++q) {
            int i = order[q];
            cd r = 0;
            for (int j = 0; j < m; ++j) r += Hat(m, j) * Y[(size_t)j * m + i];
            if (std::abs(r) <= tol * std::abs(theta[i])) ++nconv;
        }
        if (nconv >= nev || restarts >= maxiter) break;
        ++restarts;
        // thick restart: orthonormal basis of the `keep` wanted Ritz vectors
        std::vector<cd> Qk((size_t)m * keep);
        for (int q = 0; q < keep; ++q)
            for (int i = 0; i < m; ++i) Qk[(size_t)i * keep + q] = Y[(size_t)i * m + order[q]];
        for (int q = 0; q < keep; ++q) {
            for (int pass = 0; pass < 2; ++pass)
                for (int p = 0; p < q; ++p) {
                    cd s = 0;
                    for (int i = 0; i < m; ++i) s += std::conj(Qk[(size_t)i * keep + p]) * Qk[(size_t)i * keep + q];
                    for (int i = 0; i < m; ++i) Qk[(size_t)i * keep + q] -= s * Qk[(size_t)i * keep + p];
                }
            double nq = 0;
            for (int i = 0; i < m; ++i) nq += std::norm(Qk[(size_t)i * keep + q]);
            nq = std::sqrt(nq);
            for (int i = 0; i < m; ++i) Qk[(size_t)i * keep + q] /= nq;
        }
        // V[:, 0:keep] = V_m Qk ; v_keep = v_m
        for (int q = 0; q < keep; ++q) {
            std::vector<cd> col(m);
            for (int i = 0; i < m; ++i) col[i] = Qk[(size_t)i * keep + q];
            launch_zero(h->stream, Wb + (long)q * N, N);
            if ((rc = rt_update(h, Wb + (long)q * N, Vb, N, m, col.data(), 1.0))) { cleanup(); return rc; }
        }
        if (hipMemcpyAsync(vec(keep), vec(m), sizeof(double2) * N, hipMemcpyDeviceToDevice, h->stream) !=
                hipSuccess ||
            hipMemcpyAsync(Vb, Wb, sizeof(double2) * N * keep, hipMemcpyDeviceToDevice, h->stream) !=
                hipSuccess) {
            cleanup();
            return fail(h, FDFD_ERR_HIP, "restart copy failed");
        }
        // H_new = Qk^H H_m Qk, row keep = H[m, :] Qk
        std::vector<cd> HQ((size_t)m * keep, 0.0), Hn((size_t)keep * keep, 0.0), hr(keep, 0.0);
        for (int i = 0; i < m; ++i)
            for (int q = 0; q < keep; ++q)
                for (int j = 0; j < m; ++j) HQ[(size_t)i * keep + q] += Hat(i, j) * Qk[(size_t)j * keep + q];
        for (int p = 0; p < keep; ++p)
            for (int q = 0; q < keep; ++q)
                for (int i = 0; i < m; ++i)
                    Hn[(size_t)p * keep + q] += std::conj(Qk[(size_t)i * keep + p]) * HQ[(size_t)i * keep + q];
        for (int q = 0; q < keep; ++q)
            for (int j = 0; j < m; ++j) hr[q] += Hat(m, j) * Qk[(size_t)j * keep + q];
        std::fill(H.begin(), H.end(), cd(0));
        for (int p = 0; p < keep; ++p)
            for (int q = 0; q < keep; ++q) Hat(p, q) = Hn[(size_t)p * keep + q];
        for (int q = 0; q < keep; ++q) Hat(keep, q) = hr[q];
        start = keep;
    }
    // Ritz vectors of the nev wanted values, unit 2-norm, largest entry real > 0
    std::vector<cd> lam(nev), om(nev);
    std::vector<int> sel(order.begin(), order.begin() + nev);
    for (int q = 0; q < nev; ++q) {
        cd th = theta[sel[q]];
        lam[q] = sigma + 1.0 / th;
        om[q] = (pol == FDFD_TM) ? std::sqrt(-lam[q] / mu0n / eps0n) : std::sqrt(-lam[q] / mu0n);
        omegas[q].re = om[q].real();
        omegas[q].im = om[q].imag();
    }
    if (fields) {
        double2* x = h->xvec;
        double2* outd = nullptr;
        if (hipMalloc(&outd, sizeof(double2) * 3 * N) != hipSuccess) {
            cleanup();
            return fail(h, FDFD_ERR_OOM, "field buffer allocation failed");
        }
        std::vector<double2> hx(N);
        for (int q = 0; q < nev; ++q) {
            std::vector<cd> col(m);
            for (int i = 0; i < m; ++i) col[i] = Y[(size_t)i * m + sel[q]];
            launch_zero(h->stream, x, N);
            if ((rc = rt_update(h, x, Vb, N, m, col.data(), 1.0))) { (void)hipFree(outd); cleanup(); return rc; }
            if ((rc = stream_copy(h, hx.data(), x, sizeof(double2) * N, hipMemcpyDeviceToHost))) {
                (void)hipFree(outd);
                cleanup();
                return rc;
            }
            double nx = 0, best = -1;
            cd ph = 1;
            for (long i = 0; i < N; ++i) {
                double a2 = hx[i].x * hx[i].x + hx[i].y * hx[i].y;
                nx += a2;
                if (a2 > best) { best = a2; ph = cd(hx[i].x, hx[i].y); }
            }
            cd scale = std::conj(ph / std::abs(ph)) / std::sqrt(nx);
            launch_scale_copy(h->stream, x, x, make_double2(scale.real(), scale.imag()), N);
            const double c = (pol == FDFD_TM) ? mu0n : eps0n;
            cd w = 1.0 / (cd(0, 1) * om[q] * c);
            launch_fields(h->stream, outd, x, h->eps, F->sxf, F->sxb, F->syf, F->syb, g, pol,
                          make_double2(w.real(), w.imag()), 1);
            if ((rc = stream_copy(h, fields + (size_t)q * 3 * N, outd, sizeof(double2) * 3 * N,
                                  hipMemcpyDeviceToHost))) {
                (void)hipFree(outd);
                cleanup();
                return rc;
            }
        }
        (void)hipFree(outd);
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iterations = iters;
        stats->restarts = restarts;
        stats->t_factor_ms = F->t_factor_ms;
        stats->pivot_growth = F->pivot_growth;
        stats->singular_line = -1;
        stats->nconv = nconv;
    }
    cleanup();
    if (nconv < nev)
        return fail(h, FDFD_ERR_NOCONV, "Arnoldi: " + std::to_string(nconv) + " of " +
                                            std::to_string(nev) + " eigenpairs converged");
    return FDFD_OK;
}
