// fdfd_modulated_solve: the MF-FDFD driven solve of a ModulatedDevice at one
// carrier frequency, solve(d::ModulatedDevice) of src/solver/modulation.jl:35-119.
//
// The reference assembles one (2K+1) N system (modulation.jl:74-101):
//   A = blockdiag(A_{-K} .. A_K) + Cp + Cm,
//   A_j = Sxb dxb (1/mu0) Sxf dxf + Syb dyb (1/mu0) Syf dyf + wn_j^2 T_eps,
//   wn_j = w + j Omega, stretch factors at w (shared PML) or at wn_j,
//   block (j, j+1) += 0.5 wn_j^2 conj(T_deps), block (j+1, j) += 0.5 wn_{j+1}^2 T_deps,
// with b = i w src in the centre block, and solves it with lu(A)\b.
//
// Here every sideband operator A_j is factored by the block cyclic reduction
// (all 2K+1 at once, batched), and the coupled system is solved by FP64
// flexible GMRES on the exact coupled operator (stencils + the diagonal
// sideband coupling, k_mf_couple), right-preconditioned by the block
// diagonal of the factors.  The coupling is diagonal and usually weak next to
// the sideband operators, so a handful of iterations reach 1e-10.  Fields per
// sideband use the forward differences of modulation.jl:108-109.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.h"

using namespace fdfd;

namespace {
constexpr int MF_MAX_SIDEBANDS = 16;
constexpr int MF_MAX_RESTARTS = 40;
}

extern "C" int fdfd_modulated_solve(fdfd_handle* hh, double omega, double Omega, int nsidebands,
                                    int sharedpml, const fdfd_complex* deps,
                                    const fdfd_complex* src, int precision, double tol,
                                    fdfd_complex* out, fdfd_stats* stats) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h) return FDFD_ERR_INVALID;
    if (!deps || !src || !out) return fail(h, FDFD_ERR_INVALID, "deps, src and out are required");
    if (!h->has_grid || !h->has_eps)
        return fail(h, FDFD_ERR_STATE, "grid and eps must be set before fdfd_modulated_solve");
    if (nsidebands < 0 || nsidebands > MF_MAX_SIDEBANDS)
        return fail(h, FDFD_ERR_INVALID, "nsidebands must be in [0, " + std::to_string(MF_MAX_SIDEBANDS) + "]");
    if (!(omega > 0)) return fail(h, FDFD_ERR_INVALID, "omega must be > 0");
    if (!(omega - nsidebands * std::fabs(Omega) > 0))
        return fail(h, FDFD_ERR_INVALID, "every sideband frequency omega + j*Omega must be > 0");
    if (precision != FDFD_FP64 && precision != FDFD_FP32_IR)
        return fail(h, FDFD_ERR_INVALID, "precision must be FDFD_FP64 or FDFD_FP32_IR");
    if (!(tol > 0)) return fail(h, FDFD_ERR_INVALID, "tol must be > 0");
    (void)hipSetDevice(h->device);
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const int K = nsidebands, nf = 2 * K + 1;
    const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
    std::vector<double> wn(nf);
    for (int j = 0; j < nf; ++j) wn[j] = omega + Omega * (j - K);

    std::vector<Factor*> Fs(nf, nullptr);
    double2* work = nullptr;
    auto cleanup = [&]() {
        for (Factor* F : Fs)
            if (F) rt_free_factor(h, F);
        if (work) (void)hipFree(work);
    };
    for (int j = 0; j < nf; ++j) {
        Factor* F = new Factor();
        F->pol = FDFD_TM;
        F->omega = wn[j];
        F->omega_pml = sharedpml ? omega : wn[j];
        F->form = 1;
        F->precision = precision;
        Fs[j] = F;
    }
    int rc = rt_build_factors(h, Fs);
    if (rc) { cleanup(); return rc; }
    const double t_factor = Fs[0]->t_factor_ms;  // the batched build's device time
    const double scale = Fs[0]->scale;  // mu0' dx^2, the same for every sideband

    // device workspace: deps | b | x | Krylov V, Z | field buffer | coupling weights
    const long NF = (long)nf * N;
    const size_t n2 = (size_t)N + 2 * NF + (size_t)(GMRES_KDIM + 1) * NF + (size_t)GMRES_KDIM * NF +
                      (size_t)3 * N;
    if (hipMalloc(&work, n2 * sizeof(double2) + nf * sizeof(double)) != hipSuccess) {
        cleanup();
        return fail(h, FDFD_ERR_OOM, "modulated-solve workspace allocation failed");
    }
    double2* d_deps = work;
    double2* b = d_deps + N;
    double2* x = b + NF;
    double2* V = x + NF;
    double2* Zk = V + (size_t)(GMRES_KDIM + 1) * NF;
    double2* fld = Zk + (size_t)GMRES_KDIM * NF;
    double* cw = (double*)(fld + 3 * N);
    std::vector<double> ch(nf);
    for (int j = 0; j < nf; ++j) ch[j] = scale * 0.5 * wn[j] * wn[j] * eps0n;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return fail(h, FDFD_ERR_HIP, "event creation failed");
    }
    auto done = [&](int code) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        cleanup();
        return code;
    };
    (void)hipEventRecord(e0, h->stream);
    if ((rc = stream_copy(h, d_deps, deps, sizeof(double2) * N, hipMemcpyHostToDevice))) return done(rc);
    if ((rc = stream_copy(h, cw, ch.data(), sizeof(double) * nf, hipMemcpyHostToDevice))) return done(rc);
    // b' = scale * i w src in the centre block (modulation.jl:67-69), zero elsewhere
    launch_zero(h->stream, b, NF);
    if ((rc = stream_copy(h, fld, src, sizeof(double2) * N, hipMemcpyHostToDevice))) return done(rc);
    launch_scale_copy(h->stream, b + (long)K * N, fld, make_double2(0.0, scale * omega), N);
    launch_zero(h->stream, x, NF);

    auto apply = [&](const double2* xx, double2* yy, const double2* bb) -> int {
        for (int j = 0; j < nf; ++j)
            rt_apply(h, Fs[j], yy + (long)j * N, xx + (long)j * N, 0, bb ? bb + (long)j * N : nullptr);
        if (K > 0) launch_mf_couple(h->stream, yy, xx, d_deps, cw, nf, N, bb ? -1.0 : 1.0);
        return FDFD_OK;
    };
    auto prec = [&](const double2* v, double2* z) -> int {
        for (int j = 0; j < nf; ++j) {
            const int r = rt_prec_apply(h, Fs[j], 0, v + (long)j * N, z + (long)j * N);
            if (r) return r;
        }
        return FDFD_OK;
    };
    int iters = 0, restarts = 0;
    double resid = 0;
    rc = rt_gmres_op(h, NF, apply, prec, V, Zk, b, x, tol, MF_MAX_RESTARTS, &iters, &restarts, &resid);
    if (rc == FDFD_OK) {
        // fields per sideband: Hx = -Syf dyf ez / (i wn mu0'), Hy = Sxf dxf ez / (i wn mu0')
        for (int j = 0; j < nf && rc == FDFD_OK; ++j) {
            const double2 w = make_double2(0.0, -1.0 / (wn[j] * mu0n));
            launch_fields(h->stream, fld, x + (long)j * N, h->eps, Fs[j]->sxf, Fs[j]->sxb,
                          Fs[j]->syf, Fs[j]->syb, g, FDFD_TM, w, 1);
            rc = stream_copy(h, out + (size_t)j * 3 * N, fld, sizeof(double2) * 3 * N,
                             hipMemcpyDeviceToHost);
        }
    }
    (void)hipEventRecord(e1, h->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = iters;
        stats->restarts = restarts;
        stats->residual = resid;
        stats->t_factor_ms = t_factor;
        stats->t_solve_ms = ms;
        double growth = 0;
        for (Factor* F : Fs) growth = std::max(growth, F->pivot_growth);
        stats->pivot_growth = growth;
        stats->singular_line = -1;
        for (Factor* F : Fs) stats->pivot_fallback |= F->pivot_fallback;
    }
    if (rc == FDFD_OK && hipGetLastError() != hipSuccess)
        rc = fail(h, FDFD_ERR_HIP, "kernel launch failed");
    return done(rc);
}
