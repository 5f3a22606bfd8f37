// Host-side AddressSanitizer / UndefinedBehaviorSanitizer run of libfdfd.
//
// Built by build_lib.build_asan() from the library's sources with the host
// code instrumented (-Xarch_host -fsanitize=address,undefined; device code is
// not), linked into this program, and run on a GPU by
// tests/test_gpu_asan.py.  It drives every host path of the C ABI on small
// problems -- grid and eps set-up, FP64 and mixed factorizations, single and
// multi-source solves, transposed solves, the solve pullback, batched runs
// over uneven grids, eigenfrequencies, 1-D modes, full-vector modes, the
// modulated solve, Poynting and flux (with and without allocation), the
// autotuner, failure injection and argument errors -- so that a heap
// overflow, use-after-free or undefined behaviour in the host runtime
// (job tables, plans, memory pools, batch scheduling, FGMRES bookkeeping)
// aborts the run.  Numerical parity is covered by the other GPU tests; here
// only the status codes and finite results are checked.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fdfd.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what, fdfd_handle* h = nullptr) {
    if (ok) {
        printf("OK %s\n", what);
    } else {
        printf("FAIL %s: %s\n", what, h ? fdfd_last_error(h) : "");
        ++failures;
    }
    fflush(stdout);
}

bool finite(const std::vector<fdfd_complex>& v) {
    for (const auto& z : v)
        if (!std::isfinite(z.re) || !std::isfinite(z.im)) return false;
    return true;
}

constexpr double C0 = 2.99792458e14;  // um/s (close enough for frequencies here)

}  // namespace

int main() {
    fdfd_handle* h = nullptr;
    expect(fdfd_create(0, &h) == FDFD_OK, "create");
    if (!h) return 1;
    const int Nx = 72, Ny = 56;
    const long N = (long)Nx * Ny;
    const double dx = 0.02, om = 2 * M_PI * C0 / 1.55;
    std::vector<fdfd_complex> eps(N, fdfd_complex{1, 0}), src(N, fdfd_complex{0, 0});
    for (int j = 20; j < 36; ++j)
        for (int i = 0; i < Nx; ++i) eps[(long)j * Nx + i] = fdfd_complex{12.25, 0};
    for (int i = 30; i < 40; ++i)
        for (int j = 10; j < 16; ++j) eps[(long)j * Nx + i] = fdfd_complex{6.0, 0.1};
    src[(long)28 * Nx + 12] = fdfd_complex{0, 1};

    // argument errors before any device work
    expect(fdfd_set_grid(h, 0, 4, dx, dx, 0, 0, 1e-6) == FDFD_ERR_INVALID, "set_grid rejects Nx = 0", h);
    int fid = -1;
    fdfd_stats st{};
    expect(fdfd_factor(h, 1, om, 0, &fid, &st) == FDFD_ERR_STATE, "factor before grid", h);
    expect(fdfd_set_grid(h, Nx, Ny, dx, dx, 8, 8, 1e-6) == FDFD_OK, "set_grid", h);
    expect(fdfd_set_eps(h, eps.data()) == FDFD_OK, "set_eps", h);
    expect(fdfd_factor(h, 3, om, 0, &fid, &st) == FDFD_ERR_INVALID, "factor rejects pol 3", h);
    expect(fdfd_release_factor(h, 12345) == FDFD_ERR_NOTFOUND, "release unknown factor", h);

    // factor + solves, both polarizations and precisions
    for (int pol = 1; pol <= 2; ++pol)
        for (int prec = 0; prec <= 1; ++prec) {
            char name[64];
            snprintf(name, sizeof name, "factor pol %d precision %d", pol, prec);
            expect(fdfd_factor(h, pol, om, prec, &fid, &st) == FDFD_OK, name, h);
            std::vector<fdfd_complex> out(3 * N), srcs(5 * N), outs(15 * N), x(N);
            expect(fdfd_solve(h, fid, 1, src.data(), out.data(), 1e-12, &st) == FDFD_OK && finite(out),
                   "solve", h);
            for (int r = 0; r < 5; ++r) srcs[(long)r * N + 10 + 7 * r + (long)(20 + r) * Nx] = fdfd_complex{0, 1};
            expect(fdfd_solve(h, fid, 5, srcs.data(), outs.data(), 1e-12, &st) == FDFD_OK && finite(outs),
                   "solve 5 sources", h);
            expect(fdfd_linsolve(h, fid, 1, 1, src.data(), x.data(), 1e-12, &st) == FDFD_OK && finite(x),
                   "transposed linsolve", h);
            std::vector<fdfd_complex> y(N);
            double rel = 0;
            expect(fdfd_apply(h, fid, 0, x.data(), y.data()) == FDFD_OK, "apply", h);
            expect(fdfd_residual(h, fid, x.data(), src.data(), &rel) == FDFD_OK, "residual", h);
            if (pol == 1) {
                std::vector<fdfd_complex> ez(out.begin(), out.begin() + N), bar(3 * N, fdfd_complex{0, 0});
                for (long k = 0; k < N; k += 97) bar[k] = fdfd_complex{1, -1};
                for (long k = N; k < 3 * N; k += 131) bar[k] = fdfd_complex{0.5, 0.25};
                std::vector<double> g(N);
                expect(fdfd_solve_pullback(h, fid, ez.data(), bar.data(), 1e-12, g.data(), &st) == FDFD_OK,
                       "solve pullback", h);
            }
            expect(fdfd_release_factor(h, fid) == FDFD_OK, "release", h);
        }

    // failure injection
    expect(fdfd_debug_inject(h, FDFD_INJECT_SINGULAR, 3) == FDFD_OK, "inject singular");
    expect(fdfd_factor(h, 1, om, 1, &fid, &st) == FDFD_ERR_SINGULAR, "injected zero pivot", h);
    expect(fdfd_debug_inject(h, FDFD_INJECT_OOM, 0) == FDFD_OK, "inject oom");
    expect(fdfd_factor(h, 1, om, 1, &fid, &st) == FDFD_ERR_OOM, "injected allocation failure", h);
    expect(fdfd_debug_inject(h, FDFD_INJECT_PIVOT_FALLBACK, 0) == FDFD_OK, "inject fallback");
    expect(fdfd_factor(h, 1, om, 1, &fid, &st) == FDFD_OK && st.pivot_fallback == 1, "pivot fallback", h);
    {
        std::vector<fdfd_complex> out(3 * N);
        expect(fdfd_debug_inject(h, FDFD_INJECT_NOCONV, 1) == FDFD_OK, "inject noconv");
        expect(fdfd_solve(h, fid, 1, src.data(), out.data(), 1e-13, &st) == FDFD_ERR_NOCONV,
               "injected refinement failure", h);
        expect(fdfd_release_factor(h, fid) == FDFD_OK, "release", h);
    }

    // eigenfrequencies (LM and LR), fields
    {
        std::vector<fdfd_complex> w(3), fields(3 * 3 * N);
        expect(fdfd_eigenfrequency(h, 1, om, 3, 0, FDFD_WHICH_LM, 1, 1e-10, 300, w.data(), fields.data(),
                                   &st) == FDFD_OK && finite(w), "eigenfrequency LM", h);
        expect(fdfd_eigenfrequency(h, 2, om, 3, 0, FDFD_WHICH_LR, 0, 1e-10, 300, w.data(), nullptr,
                                   &st) == FDFD_OK && finite(w), "eigenfrequency LR", h);
        expect(fdfd_eigenfrequency(h, 1, om, 3, 0, 9, 0, 1e-10, 300, w.data(), nullptr, &st) ==
                   FDFD_ERR_INVALID, "eigenfrequency rejects which", h);
    }

    // 1-D modes (two slices), full-vector modes
    {
        std::vector<fdfd_complex> s1(41, fdfd_complex{1, 0}), s2(51, fdfd_complex{1, 0});
        for (int i = 15; i < 26; ++i) s1[i] = fdfd_complex{12.25, 0};
        for (int i = 20; i < 31; ++i) s2[i] = fdfd_complex{12.25, 0};
        fdfd_mode_problem pr[2] = {{41, 1, om, 0.02, {3.0, 0}, s1.data()},
                                   {51, 2, om, 0.02, {3.0, 0}, s2.data()}};
        std::vector<fdfd_complex> beta(2), vecs(41 + 51);
        int nconv[2] = {0, 0};
        expect(fdfd_eigenmode(h, 2, pr, 1, 0, 1e-6, 0, 0, beta.data(), vecs.data(), nconv, &st) == FDFD_OK,
               "eigenmode batch", h);
        std::vector<fdfd_complex> vb(1), vf(6 * N);
        expect(fdfd_vector_eigenmode(h, om, fdfd_complex{2.5, 0}, 1, 0, 1, 1e-9, 300, vb.data(), vf.data(),
                                     &st) == FDFD_OK && finite(vb), "vector eigenmode", h);
    }

    // modulated solve (K = 1)
    {
        std::vector<fdfd_complex> deps(N, fdfd_complex{0, 0}), out(3 * 3 * N);
        for (int j = 20; j < 28; ++j)
            for (int i = 20; i < 50; ++i) deps[(long)j * Nx + i] = fdfd_complex{std::cos(0.3 * i), std::sin(0.3 * i)};
        expect(fdfd_modulated_solve(h, om, 0.02 * om, 1, 1, deps.data(), src.data(), 1, 1e-10, out.data(),
                                    &st) == FDFD_OK && finite(out), "modulated solve", h);
    }

    // Poynting and flux, host and device-pointer and no-allocation entries
    {
        std::vector<fdfd_complex> out(3 * N);
        expect(fdfd_factor(h, 1, om, 0, &fid, &st) == FDFD_OK, "factor", h);
        expect(fdfd_solve(h, fid, 1, src.data(), out.data(), 1e-12, &st) == FDFD_OK, "solve", h);
        (void)fdfd_release_factor(h, fid);
        std::vector<double> S(2 * N), fl(3), fl2(3);
        fdfd_flux_plane pl[3] = {{1, 40, 0, Ny - 1}, {2, 30, 5, 60}, {1, Nx - 2, 10, 20}};
        expect(fdfd_poynting(h, 1, Nx, Ny, out.data(), S.data()) == FDFD_OK, "poynting", h);
        expect(fdfd_flux(h, 1, Nx, Ny, dx, dx, out.data(), 3, pl, fl.data()) == FDFD_OK, "flux", h);
        fdfd_flux_plane bad = {1, Nx - 1, 0, 3};
        expect(fdfd_flux(h, 1, Nx, Ny, dx, dx, out.data(), 1, &bad, fl.data()) == FDFD_ERR_INVALID,
               "flux rejects a plane on the last column", h);
    }

    // batched runs over uneven grids (several sources, several frequencies)
    {
        fdfd_batch* b = nullptr;
        expect(fdfd_batch_create(h, &b) == FDFD_OK, "batch create", h);
        const int gx[2] = {48, 40}, gy[2] = {40, 64};
        const double oms[3] = {om * 0.97, om, om * 1.03};
        int ids[2];
        for (int q = 0; q < 2; ++q) {
            const long M = (long)gx[q] * gy[q];
            std::vector<fdfd_complex> e(M, fdfd_complex{2.0, 0}), s(3 * M, fdfd_complex{0, 0});
            for (int r = 0; r < 3; ++r) s[(long)r * M + (long)(10 + r) * gx[q] + 9 + 2 * r] = fdfd_complex{0, 1};
            expect(fdfd_batch_add(b, gx[q], gy[q], dx, dx, 6, 6, 1e-6, 1, e.data(), 3, oms, 3, s.data(),
                                  &ids[q]) == FDFD_OK, "batch add", h);
        }
        for (int prec = 0; prec <= 1; ++prec)
            expect(fdfd_batch_run(b, prec, 1e-11, &st) == FDFD_OK, "batch run", h);
        std::vector<fdfd_complex> f(3 * 40 * 64);
        expect(fdfd_batch_result(b, ids[1], 2, 1, f.data()) == FDFD_OK, "batch result", h);
        expect(fdfd_batch_result(b, ids[1], 3, 0, f.data()) == FDFD_ERR_INVALID, "batch result range", h);
        expect(fdfd_batch_destroy(b) == FDFD_OK, "batch destroy", h);
    }

    // autotune on a small grid (no cache file)
    {
        expect(fdfd_set_grid(h, 64, 48, dx, dx, 6, 6, 1e-6) == FDFD_OK, "set_grid", h);
        std::vector<fdfd_complex> e(64 * 48, fdfd_complex{3.0, 0});
        expect(fdfd_set_eps(h, e.data()) == FDFD_OK, "set_eps", h);
        expect(fdfd_autotune(h, 1, om, 1, &st) == FDFD_OK, "autotune mixed", h);
        int knobs[5];
        expect(fdfd_get_tuning(h, knobs, 5) == FDFD_OK, "get tuning", h);
    }

    expect(fdfd_destroy(h) == FDFD_OK, "destroy");
    printf("%s: %d failures\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 2 : 0;
}
