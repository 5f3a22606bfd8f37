// fdfd_vector_eigenmode: full-vector modes of a waveguide cross-section,
// eigenmode(d::AbstractDevice{2}, neff, nev) of src/solver/eigen.jl:31-66.
//
// The reference builds the 2N x 2N operator on (Hx, Hy) (eigen.jl:45) and
// calls eigs(A, nev, sigma = (w/c0 neff)^2).  Here the shift-invert Arnoldi
// (arnoldi.cpp) applies OP = (A - sigma)^-1 by FP64 flexible GMRES on the
// exact coupled operator: the two diagonal blocks A_xx - sigma and
// A_yy - sigma are 5-point periodic operators, factored by the block cyclic
// reduction, and preconditioned blockwise; the polarization coupling A_xy,
// A_yx (zero where eps is uniform) is applied by k_vmode_offdiag.  Fields
// follow eigen.jl:54-61.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.h"

namespace fdfd {
int rt_arnoldi(Handle* h, long N, int nev, int ncv, int which, double tol, int maxiter,
               const std::function<int(const double2*, double2*)>& op, double2* ritz,
               std::vector<cd>& theta_out, int* nconv_out, int* restarts_out,
               const double* op_error);
}

using namespace fdfd;

namespace {
constexpr int VM_MAX_RESTARTS = 40;
}

extern "C" int fdfd_vector_eigenmode(fdfd_handle* hh, double omega, fdfd_complex neff, int nev,
                                     int ncv, int precision, double tol, int maxiter,
                                     fdfd_complex* beta, fdfd_complex* fields,
                                     fdfd_stats* stats) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !beta || nev < 1) return FDFD_ERR_INVALID;
    if (!h->has_grid || !h->has_eps)
        return fail(h, FDFD_ERR_STATE, "grid and eps must be set before fdfd_vector_eigenmode");
    if (!(omega > 0)) return fail(h, FDFD_ERR_INVALID, "omega must be > 0");
    if (precision != FDFD_FP64 && precision != FDFD_FP32_IR)
        return fail(h, FDFD_ERR_INVALID, "precision must be FDFD_FP64 or FDFD_FP32_IR");
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny, N2 = 2 * N;
    if (ncv <= 0) ncv = std::max(20, 2 * nev + 1);  // Arpack.jl default
    ncv = (int)std::min<long>(ncv, N2);
    if (nev >= ncv) return fail(h, FDFD_ERR_INVALID, "nev must be < ncv (and < 2 Nx Ny)");
    if (ncv > 64) return fail(h, FDFD_ERR_INVALID, "ncv must be <= 64");
    if (!(tol > 0)) tol = 1e-12;
    if (maxiter <= 0) maxiter = 300;
    (void)hipSetDevice(h->device);
    const double eps0n = EPS0 * g.L0;
    const double c0n = std::sqrt(1.0 / EPS0 / MU0) / g.L0;
    const cd kb = omega / c0n * cd(neff.re, neff.im);
    const cd sigma = kb * kb;

    std::vector<Factor*> Fs(2, nullptr);
    double2* work = nullptr;
    auto cleanup = [&]() {
        for (Factor* F : Fs)
            if (F) rt_free_factor(h, F);
        if (work) (void)hipFree(work);
    };
    for (int b = 0; b < 2; ++b) {
        Factor* F = new Factor();
        F->pol = FDFD_TM;
        F->omega = omega;
        F->form = 2 + b;
        F->shift = make_double2(sigma.real(), sigma.imag());
        F->precision = precision;
        Fs[b] = F;
    }
    int rc = rt_build_factors(h, Fs);
    if (rc) { cleanup(); return rc; }
    const double s = Fs[0]->scale;  // dx^2
    // inner FGMRES workspace (2N vectors): V, Zk, b; Ritz vectors nev x 2N; fields 6N
    const size_t n2 = (size_t)(GMRES_KDIM + 1) * N2 + (size_t)GMRES_KDIM * N2 + N2 +
                      (size_t)nev * N2 + 6 * (size_t)N;
    if (hipMalloc(&work, n2 * sizeof(double2)) != hipSuccess) {
        cleanup();
        return fail(h, FDFD_ERR_OOM, "vector-mode workspace allocation failed");
    }
    double2* V = work;
    double2* Zk = V + (size_t)(GMRES_KDIM + 1) * N2;
    double2* bb = Zk + (size_t)GMRES_KDIM * N2;
    double2* ritz = bb + N2;
    double2* fld = ritz + (size_t)nev * N2;

    // s (A - sigma) on (Hx, Hy): the factored blocks' planes plus the coupling
    auto apply = [&](const double2* x, double2* y, const double2* b) -> int {
        launch_apply(h->stream, y, x, Fs[0]->coef, g.Nx, g.Ny, 0, b);
        launch_apply(h->stream, y + N, x + N, Fs[1]->coef, g.Nx, g.Ny, 0, b ? b + N : nullptr);
        launch_vmode_offdiag(h->stream, y, x, h->eps, g, b ? -s : s);
        return FDFD_OK;
    };
    auto prec = [&](const double2* v, double2* z) -> int {
        int r = rt_prec_apply(h, Fs[0], 0, v, z);
        if (!r) r = rt_prec_apply(h, Fs[1], 0, v + N, z + N);
        return r;
    };
    const double inner_tol = std::max(1e-13, std::min(1e-12, tol * 1e-1));
    int inner_its = 0, inner_restarts = 0;
    double op_err = 0;  // largest relative residual of the OP applications
    // OP y = (A - sigma)^-1 x: solve s (A - sigma) y = s x.  With the shift
    // very close to an eigenvalue, A - sigma is nearly singular and FP64
    // rounding floors the residual above inner_tol (about |A| eps / |l - sigma|);
    // a solve that stagnates below 1e-8 is accepted and Arnoldi's tolerance
    // follows its error.
    auto op = [&](const double2* x, double2* y) -> int {
        launch_scale_copy(h->stream, bb, x, make_double2(s, 0), N2);
        launch_zero(h->stream, y, N2);
        double res = 0;
        int r = rt_gmres_op(h, N2, apply, prec, V, Zk, bb, y, inner_tol, VM_MAX_RESTARTS, &inner_its,
                            &inner_restarts, &res);
        if (r == FDFD_ERR_NOCONV && res <= 1e-8) r = FDFD_OK;
        op_err = std::max(op_err, res);
        return r;
    };
    std::vector<cd> theta;
    int nconv = 0, restarts = 0;
    rc = rt_arnoldi(h, N2, nev, ncv, FDFD_WHICH_LM, tol, maxiter, op, fields ? ritz : nullptr, theta,
                    &nconv, &restarts, &op_err);
    if (rc) { cleanup(); return rc; }
    std::vector<cd> be(nev);
    for (int q = 0; q < nev; ++q) {
        be[q] = std::sqrt(sigma + 1.0 / theta[q]);  // eigen.jl:51, principal branch
        beta[q].re = be[q].real();
        beta[q].im = be[q].imag();
    }
    if (fields) {
        for (int q = 0; q < nev && rc == FDFD_OK; ++q) {
            launch_vmode_fields(h->stream, fld, ritz + (size_t)q * N2, h->eps, g,
                                make_double2(be[q].real(), be[q].imag()), omega, eps0n);
            rc = stream_copy(h, fields + (size_t)q * 6 * N, fld, sizeof(double2) * 6 * N,
                             hipMemcpyDeviceToHost);
        }
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = inner_its;
        stats->restarts = restarts;
        stats->t_factor_ms = Fs[0]->t_factor_ms;
        stats->pivot_growth = std::max(Fs[0]->pivot_growth, Fs[1]->pivot_growth);
        stats->singular_line = -1;
        stats->nconv = nconv;
        stats->pivot_fallback = Fs[0]->pivot_fallback | Fs[1]->pivot_fallback;
    }
    cleanup();
    if (rc) return rc;
    if (nconv < nev)
        return fail(h, FDFD_ERR_NOCONV, "Arnoldi: " + std::to_string(nconv) + " of " +
                                            std::to_string(nev) + " eigenpairs converged");
    return FDFD_OK;
}
