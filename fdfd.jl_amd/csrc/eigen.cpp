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
// device (FP64, CGS2 orthogonalisation) in the shared thick-restart
// Arnoldi routine (arnoldi.cpp).
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
int rt_arnoldi(Handle* h, long N, int nev, int ncv, int which, double tol, int maxiter,
               const std::function<int(const double2*, double2*)>& op, double2* ritz,
               std::vector<cd>& theta_out, int* nconv_out, int* restarts_out,
               const double* op_error);

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
    if (precision != FDFD_FP64 && precision != FDFD_FP32_IR)
        return fail(h, FDFD_ERR_INVALID, "precision must be FDFD_FP64 or FDFD_FP32_IR");
    if (which != FDFD_WHICH_LM && which != FDFD_WHICH_LR)
        return fail(h, FDFD_ERR_INVALID, "which must be FDFD_WHICH_LM or FDFD_WHICH_LR");
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

    int iters = 0;
    double op_err = 0;  // largest relative residual of the OP applications
    // OP y = x.  With an eigenvalue very close to the shift, A - sigma is nearly
    // singular and FP64 rounding floors the solve's residual above inner_tol
    // (about |A| eps / |lambda - sigma|); a solve that stagnates below 1e-8 is
    // accepted and Arnoldi's convergence test follows its error.
    auto op = [&](const double2* x, double2* y) -> int {
        if (pol == FDFD_TM) launch_mul_eps(h->stream, h->tmp, x, h->eps, make_double2(opscale, 0), N);
        else launch_scale_copy(h->stream, h->tmp, x, make_double2(opscale, 0), N);
        launch_zero(h->stream, y, N);
        double res = 0;
        int r = rt_gmres(h, F, 0, h->tmp, y, inner_tol, &iters, &res);
        if (r == FDFD_ERR_NOCONV && res <= 1e-8) r = FDFD_OK;
        op_err = std::max(op_err, res);
        return r;
    };
    double2* ritz = nullptr;
    if (fields && hipMalloc(&ritz, sizeof(double2) * N * nev) != hipSuccess) {
        rt_free_factor(h, F);
        return fail(h, FDFD_ERR_OOM, "Ritz vector allocation failed");
    }
    auto cleanup = [&]() {
        if (ritz) (void)hipFree(ritz);
        rt_free_factor(h, F);
    };
    std::vector<cd> theta;
    int nconv = 0, restarts = 0;
    rc = rt_arnoldi(h, N, nev, ncv, which, tol, maxiter, op, ritz, theta, &nconv, &restarts, &op_err);
    if (rc) { cleanup(); return rc; }
    std::vector<cd> lam(nev), om(nev);
    for (int q = 0; q < nev; ++q) {
        lam[q] = sigma + 1.0 / theta[q];
        om[q] = (pol == FDFD_TM) ? std::sqrt(-lam[q] / mu0n / eps0n) : std::sqrt(-lam[q] / mu0n);
        omegas[q].re = om[q].real();
        omegas[q].im = om[q].imag();
    }
    if (fields) {
        double2* outd = nullptr;
        if (hipMalloc(&outd, sizeof(double2) * 3 * N) != hipSuccess) {
            cleanup();
            return fail(h, FDFD_ERR_OOM, "field buffer allocation failed");
        }
        for (int q = 0; q < nev; ++q) {
            const double c = (pol == FDFD_TM) ? mu0n : eps0n;
            cd w = 1.0 / (cd(0, 1) * om[q] * c);
            launch_fields(h->stream, outd, ritz + (long)q * N, h->eps, F->sxf, F->sxb, F->syf, F->syb,
                          g, pol, make_double2(w.real(), w.imag()), 1);
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
