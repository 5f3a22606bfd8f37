// Batched driven solves: many grids x frequencies x sources.
//
// The reference solves one device at a time and refactors per frequency
// (src/solver/driven.jl:11-55), each source needing its own solve() call.
// Here all frequencies of one grid that fit in HBM are factored together:
// every factorization kernel takes a job table, so B frequencies become 2B
// independent elimination chains in the same launches (B times the work per
// launch for the same number of launches).  Each (frequency, source) pair is
// then solved with FP64 FGMRES on its factor; fields stay in device memory
// until fdfd_batch_result copies one set out.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/fdfd.h"
#include "runtime.h"

namespace fdfd {
int rt_gmres(Handle* h, Factor* F, int trans, const double2* b, double2* x, double tol, int* iters,
             double* resid);
int rt_build_factors(Handle* h, const std::vector<Factor*>& Fs);
}  // namespace fdfd

using namespace fdfd;

struct BatchJob {
    GridDesc g{};
    int pol = FDFD_TM;
    std::vector<double2> eps;
    std::vector<double> omegas;
    int nsrc = 0;
    std::vector<double2> srcs;
    double2* result = nullptr;  // device: nomega * nsrc * 3N
};

struct fdfd_batch {
    fdfd_handle* hh = nullptr;
    std::vector<BatchJob> jobs;
    bool ran = false;
};

static void free_results(fdfd_batch* b) {
    for (auto& j : b->jobs)
        if (j.result) {
            (void)hipFree(j.result);
            j.result = nullptr;
        }
}

extern "C" {

int fdfd_batch_create(fdfd_handle* hh, fdfd_batch** out) {
    if (!hh || !out) return FDFD_ERR_INVALID;
    auto* b = new fdfd_batch();
    b->hh = hh;
    *out = b;
    return FDFD_OK;
}

int fdfd_batch_add(fdfd_batch* b, int Nx, int Ny, double dx, double dy, int npml_x, int npml_y,
                   double L0, int pol, const fdfd_complex* eps, int nomega, const double* omegas,
                   int nsrc, const fdfd_complex* srcs, int* job_id) {
    if (!b || !eps || !omegas || !srcs || nomega < 1 || nsrc < 1) return FDFD_ERR_INVALID;
    Handle* h = &b->hh->impl;
    if (Nx < 1 || Ny < 1 || !(dx > 0) || !(dy > 0) || npml_x < 0 || npml_y < 0 || !(L0 > 0))
        return fail(h, FDFD_ERR_INVALID, "batch job: invalid grid parameters");
    if (std::min(Nx, Ny) > MAX_LINE)
        return fail(h, FDFD_ERR_INVALID, "batch job: min(Nx, Ny) must be <= 1024");
    if (pol != FDFD_TM && pol != FDFD_TE) return fail(h, FDFD_ERR_INVALID, "batch job: bad pol");
    for (int i = 0; i < nomega; ++i)
        if (!(omegas[i] > 0)) return fail(h, FDFD_ERR_INVALID, "batch job: omega must be > 0");
    BatchJob j;
    j.g = GridDesc{Nx, Ny, dx, dy, npml_x, npml_y, L0};
    j.pol = pol;
    const long N = (long)Nx * Ny;
    j.eps.assign((const double2*)eps, (const double2*)eps + N);
    j.omegas.assign(omegas, omegas + nomega);
    j.nsrc = nsrc;
    j.srcs.assign((const double2*)srcs, (const double2*)srcs + N * nsrc);
    b->jobs.push_back(std::move(j));
    if (job_id) *job_id = (int)b->jobs.size() - 1;
    b->ran = false;
    return FDFD_OK;
}

int fdfd_batch_run(fdfd_batch* b, int precision, double tol, fdfd_stats* stats) {
    if (!b) return FDFD_ERR_INVALID;
    fdfd_handle* hh = b->hh;
    Handle* h = &hh->impl;
    if (precision != FDFD_FP64 && precision != FDFD_FP32_IR)
        return fail(h, FDFD_ERR_INVALID, "precision must be FDFD_FP64 or FDFD_FP32_IR");
    if (!(tol > 0)) return fail(h, FDFD_ERR_INVALID, "tol must be > 0");
    (void)hipSetDevice(h->device);
    const bool timing = std::getenv("FDFD_DEBUG_TIMING") != nullptr;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto secs = [](auto a, auto c) { return std::chrono::duration<double>(c - a).count(); };
    const auto tr0 = now();
    free_results(b);
    // the batch replaces the handle's problem: its factors are released and
    // it is left without a grid (fdfd_set_grid starts the next problem)
    (void)hipStreamSynchronize(h->stream);
    for (auto& kv : h->factors) rt_free_factor(h, kv.second);
    h->factors.clear();
    h->has_grid = h->has_eps = false;
    // largest factors first (the slab below is sized by them)
    std::vector<int> order(b->jobs.size());
    std::iota(order.begin(), order.end(), 0);
    auto fbytes = [&](const BatchJob& j) {
        const int n = std::min(j.g.Nx, j.g.Ny), m = std::max(j.g.Nx, j.g.Ny);
        return rt_factor_bytes(n, m, precision, h->gemm_b3 != 0);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return fbytes(b->jobs[a]) > fbytes(b->jobs[c]); });
    // Memory plan: result buffers first, then one slab for the factor arenas
    // of the largest group of frequencies, reused by every group of every
    // grid (no factor-sized allocation or free inside the batch).
    arena_clear(h);
    int rc = FDFD_OK;
    const auto ta0 = now();
    if (timing) fprintf(stderr, "batch release of the handle's factors and pool: %.3f s\n", secs(tr0, ta0));
    for (auto& job : b->jobs) {
        const size_t nres = (size_t)job.omegas.size() * job.nsrc * 3 * job.g.Nx * (size_t)job.g.Ny;
        if (hipMalloc(&job.result, sizeof(double2) * nres) != hipSuccess) {
            job.result = nullptr;
            free_results(b);
            return fail(h, FDFD_ERR_OOM, "batch: result buffer allocation failed");
        }
    }
    {
        long Nmax = 0;
        int kmax = 1;
        for (const auto& job : b->jobs) {
            Nmax = std::max(Nmax, (long)job.g.Nx * job.g.Ny);
            kmax = std::max(kmax, std::min(job.nsrc, KMULTI));
        }
        if ((rc = ensure_solve_work(h, Nmax)) || (rc = ensure_multi_work(h, Nmax, kmax))) {
            free_results(b);
            return rc;
        }
    }
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    // factor memory of a group: at most FDFD_BATCH_MEMORY_GB (default 64) GB;
    // larger groups gain little once a level's launches fill the chip
    double budget = 64.0;
    if (const char* eb = std::getenv("FDFD_BATCH_MEMORY_GB")) budget = std::max(1.0, std::atof(eb));
    const double usable = std::min(0.8 * (double)freeb, budget * 1073741824.0);
    std::vector<int> group(b->jobs.size(), 1);
    std::vector<size_t> arena_bytes(b->jobs.size(), 0);
    size_t slab_need = 0;
    for (size_t ji = 0; ji < b->jobs.size(); ++ji) {
        const BatchJob& job = b->jobs[ji];
        const int n = std::min(job.g.Nx, job.g.Ny), m = std::max(job.g.Nx, job.g.Ny);
        size_t sb = 0, ss = 0;
        rt_factor_arena_sizes(n, m, precision, h->gemm_b3 != 0, &sb, &ss);
        arena_bytes[ji] = ((sb + 255) & ~(size_t)255) + ((ss + 255) & ~(size_t)255);
        // per frequency: the arenas plus the operator, Gauss-Jordan work and
        // tables allocated beside them (about 96 B and 2 GJ channels per cell)
        const double N = (double)job.g.Nx * job.g.Ny;
        const double per = (double)arena_bytes[ji] + 96.0 * N + 64.0 * N * (double)VMAX / n;
        int B = (int)std::floor(usable / per);
        B = std::max(1, std::min<int>({B, (int)job.omegas.size(), 64}));
        group[ji] = B;
        slab_need = std::max(slab_need, (size_t)B * arena_bytes[ji]);
    }
    if (hipMalloc((void**)&h->slab, slab_need) != hipSuccess) {
        (void)hipGetLastError();
        h->slab = nullptr;  // factors then allocate their arenas one by one
        h->slab_size = 0;
    } else {
        h->slab_size = slab_need;
    }
    h->slab_used = 0;
    if (timing)
        fprintf(stderr, "batch setup (results %.1f GB, slab %.1f GB): %.3f s\n",
                [&] { double t = 0; for (auto& j : b->jobs) t += (double)j.omegas.size() * j.nsrc * 48.0 * j.g.Nx * j.g.Ny; return t / 1e9; }(),
                slab_need / 1e9, secs(ta0, now()));
    int iters = 0, pivot_fallback = 0;
    double worst = 0, t_factor = 0, t_solve = 0, growth = 0;
    for (int ji : order) {
        BatchJob& job = b->jobs[ji];
        const auto tj0 = now();
        const GridDesc& g = job.g;
        const long N = (long)g.Nx * g.Ny;
        h->g = g;
        h->has_grid = true;
        if (h->eps_n < N) {
            if (h->eps) (void)hipFree(h->eps);
            h->eps = nullptr;
            h->eps_n = 0;
            if (hipMalloc(&h->eps, sizeof(double2) * N) != hipSuccess) {
                h->eps = nullptr;
                rc = fail(h, FDFD_ERR_OOM, "batch: eps buffer allocation failed");
                break;
            }
            h->eps_n = N;
        }
        if ((rc = stream_copy(h, h->eps, job.eps.data(), sizeof(double2) * N, hipMemcpyHostToDevice)))
            break;
        h->has_eps = true;
        double2* dsrc = nullptr;
        if (hipMalloc(&dsrc, sizeof(double2) * N * job.nsrc) != hipSuccess ||
            stream_copy(h, dsrc, job.srcs.data(), sizeof(double2) * N * job.nsrc,
                        hipMemcpyHostToDevice) != FDFD_OK) {
            if (dsrc) (void)hipFree(dsrc);
            rc = fail(h, FDFD_ERR_OOM, "batch: source upload failed");
            break;
        }
        const int B = group[ji];
        double tf = 0, ts = 0;
        for (size_t o0 = 0; o0 < job.omegas.size() && rc == FDFD_OK; o0 += B) {
            const int nb = (int)std::min<size_t>(B, job.omegas.size() - o0);
            std::vector<Factor*> Fs;
            for (int q = 0; q < nb; ++q) {
                Factor* F = new Factor();
                F->pol = job.pol;
                F->omega = job.omegas[o0 + q];
                F->precision = precision;
                Fs.push_back(F);
            }
            const auto tb0 = now();
            h->slab_used = 0;  // the previous group's factors are freed
            rc = rt_build_factors(h, Fs);
            tf += secs(tb0, now());
            for (Factor* F : Fs) {
                pivot_fallback |= F->pivot_fallback;
                growth = std::max(growth, F->pivot_growth);
            }
            if (rc == FDFD_OK) {
                t_factor += Fs[0]->t_factor_ms;
                const auto ts0 = now();
                hipEvent_t e0, e1;
                (void)hipEventCreate(&e0);
                (void)hipEventCreate(&e1);
                (void)hipEventRecord(e0, h->stream);
                const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
                for (int q = 0; q < nb && rc == FDFD_OK; ++q) {
                    Factor* F = Fs[q];
                    const double c = (F->pol == FDFD_TM) ? mu0n : eps0n;
                    const double2 w = make_double2(0.0, -1.0 / (F->omega * c));
                    // the sources of this frequency, KMULTI at a time, in one
                    // lockstep FGMRES on the factor
                    for (int s0 = 0; s0 < job.nsrc && rc == FDFD_OK; s0 += KMULTI) {
                        const int ns = std::min(KMULTI, job.nsrc - s0);
                        for (int k = 0; k < ns; ++k) {
                            // b' = scale * (i omega) * src   (driven.jl:36,46)
                            launch_scale_copy(h->stream, h->mB + (long)k * N, dsrc + (long)(s0 + k) * N,
                                              make_double2(0.0, F->scale * F->omega), N);
                            launch_zero(h->stream, h->mX + (long)k * N, N);
                        }
                        rc = rt_gmres_multi(h, F, ns, h->mB, h->mX, tol, &iters, &worst);
                        if (rc) break;
                        for (int k = 0; k < ns; ++k) {
                            double2* out = job.result + ((o0 + q) * job.nsrc + s0 + k) * 3 * N;
                            launch_fields(h->stream, out, h->mX + (long)k * N, h->eps, F->sxf,
                                          F->sxb, F->syf, F->syb, g, F->pol, w, 0);
                        }
                    }
                }
                (void)hipEventRecord(e1, h->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                t_solve += ms;
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                ts += secs(ts0, now());
            }
            const auto tr0 = now();
            for (Factor* F : Fs) rt_free_factor(h, F);
            tf += secs(tr0, now());
        }
        (void)hipFree(dsrc);
        if (timing)
            fprintf(stderr, "batch job %d (%d x %d, %zu omegas, %d sources, B=%d): %.3f s "
                    "(factor incl. host %.3f s, solves %.3f s)\n", ji, g.Nx, g.Ny,
                    job.omegas.size(), job.nsrc, B, secs(tj0, now()), tf, ts);
        if (rc) break;
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->slab);
    h->slab = nullptr;
    h->slab_size = h->slab_used = 0;
    h->has_grid = h->has_eps = false;
    if (stats) {  // also on failure: what was reached
        memset(stats, 0, sizeof(*stats));
        stats->iterations = iters;
        stats->residual = worst;
        stats->t_factor_ms = t_factor;
        stats->t_solve_ms = t_solve;
        stats->pivot_growth = growth;
        stats->singular_line = -1;
        stats->pivot_fallback = pivot_fallback;
    }
    if (rc) {
        free_results(b);
        return rc;
    }
    b->ran = true;
    return FDFD_OK;
}

int fdfd_batch_result(fdfd_batch* b, int job_id, int iomega, int isrc, fdfd_complex* fields) {
    if (!b || !fields) return FDFD_ERR_INVALID;
    Handle* h = &b->hh->impl;
    if (!b->ran) return fail(h, FDFD_ERR_STATE, "fdfd_batch_run has not completed");
    if (job_id < 0 || job_id >= (int)b->jobs.size()) return fail(h, FDFD_ERR_NOTFOUND, "bad job id");
    const BatchJob& j = b->jobs[job_id];
    if (iomega < 0 || iomega >= (int)j.omegas.size() || isrc < 0 || isrc >= j.nsrc)
        return fail(h, FDFD_ERR_INVALID, "frequency or source index out of range");
    const long N = (long)j.g.Nx * j.g.Ny;
    const double2* src = j.result + ((long)iomega * j.nsrc + isrc) * 3 * N;
    return stream_copy(h, fields, src, sizeof(double2) * 3 * N, hipMemcpyDeviceToHost);
}

int fdfd_batch_destroy(fdfd_batch* b) {
    if (!b) return FDFD_ERR_INVALID;
    free_results(b);
    delete b;
    return FDFD_OK;
}

}  // extern "C"
