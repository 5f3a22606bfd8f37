// Batched driven solves: many grids x frequencies x sources.
//
// The reference solves one device at a time and refactors per frequency
// (src/solver/driven.jl:11-55), each source needing its own solve() call.
// Here the frequencies of one grid are factored in groups: every
// factorization kernel takes a job table, so B frequencies become 2B
// independent elimination chains in the same launches (B times the work per
// launch for the same number of launches).  The sources of each frequency
// are then solved together (lockstep multi-source FGMRES on its factor).
// The groups of all grids, largest grid first, run as a two-lane pipeline:
// a second handle factors group g + 1 on its own stream, from a worker
// thread, while the batch's handle solves group g, so the throughput-bound
// factorization kernels and the latency-bound solve steps share the GPU,
// across grids of different sizes as well.  Fields stay in device memory
// until fdfd_batch_result copies one set out.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/fdfd.h"
#include "runtime.h"

namespace fdfd {
int rt_gmres(Handle* h, Factor* F, int trans, const double2* b, double2* x, double tol, int* iters,
             double* resid);
int rt_build_factors(Handle* h, const std::vector<Factor*>& Fs);
}  // namespace fdfd

using namespace fdfd;

namespace {

// One factor group: frequencies o0 .. o0 + nb - 1 of job ji, factored together
struct Group {
    int ji, o0, nb;
};

// A second handle that factors the next group on its own stream (from a
// worker thread) while the batch's handle solves the current one.  It owns
// its stream, Gauss-Jordan work, eps buffer and a half of the factor slab;
// the reduced-solve vectors and Krylov work are borrowed from the solving
// handle (the factors' solve tables point at its x_out and g_mid).
struct Builder {
    Handle H;
    int job = -1;  // the job whose grid and eps H holds
    ~Builder() {
        Handle& b = H;
        if (b.stream) {
            (void)hipStreamSynchronize(b.stream);
            (void)hipStreamDestroy(b.stream);
        }
        for (void* p : {(void*)b.used, (void*)b.piv, (void*)b.win, (void*)b.verr, b.Vbuf, b.LU, b.Zb,
                        (void*)b.eps})
            if (p) (void)hipFree(p);
        arena_clear(&b);
        buf_clear(&b);
        if (b.tstage) (void)hipHostFree(b.tstage);
        if (b.tstage_ev) (void)hipEventDestroy(b.tstage_ev);
    }
};

int builder_init(Builder& B, Handle* h) {
    Handle& b = B.H;
    b.device = h->device;
    b.panel_width = h->panel_width;
    b.full_pivot_search = h->full_pivot_search;
    // a pending one-shot factorization failure applies to the batch's first
    // build (a solve failure stays with the solving handle)
    if (h->inject == FDFD_INJECT_PIVOT_FALLBACK || h->inject == FDFD_INJECT_SINGULAR ||
        h->inject == FDFD_INJECT_OOM) {
        b.inject = h->inject;
        b.inject_arg = h->inject_arg;
        h->inject = 0;
    }
    b.use_graphs = false;
    b.gemm3m = h->gemm3m;
    b.mv_mfma = h->mv_mfma;
    b.gemm_waves = h->gemm_waves;
    b.gemm_b3 = h->gemm_b3;
    b.trail_h3 = h->trail_h3;
    b.fuse_panels = h->fuse_panels;
    b.env_knobs = h->env_knobs;
    b.tuned_explicitly = h->tuned_explicitly;
    b.work_N = h->work_N;  // borrowed: sized for the batch's largest grid
    b.g_mid = h->g_mid;
    b.x_out = h->x_out;
    b.V = h->V;
    b.Zk = h->Zk;
    b.kdim = h->kdim;
    b.tmp = h->tmp;
    b.tmp2 = h->tmp2;
    b.bvec = h->bvec;
    b.xvec = h->xvec;
    b.dpart = h->dpart;
    b.rpart = h->rpart;
    b.hdev = h->hdev;
    b.nblk = h->nblk;
    HIPCHK(h, hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
    return FDFD_OK;
}

// Borrowed pointers are cleared before the builder is destroyed.
void builder_release_borrowed(Builder& B) {
    Handle& b = B.H;
    b.g_mid = b.x_out = b.V = b.Zk = b.tmp = b.tmp2 = b.bvec = b.xvec = b.hdev = nullptr;
    b.dpart = nullptr;
    b.rpart = nullptr;
    b.work_N = 0;
}

}  // namespace

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
    // Memory plan: result buffers first, then a slab of two halves for the
    // factor arenas of the largest group of frequencies (one half is solved
    // while the next group is factored into the other), reused by every group
    // of every grid (no factor-sized allocation or free inside the batch).
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
    // factor memory of a group: at most FDFD_BATCH_MEMORY_GB (default 32) GB;
    // larger groups gain little once a level's launches fill the chip (C3's
    // three grids: 2.40 s at 32 GB per group, 2.39 s at 64, 2.6-2.7 s at 16).
    // Two groups are held at once (one factored while the other is solved),
    // so the free memory is shared between two slabs.
    double budget = 32.0;
    if (const char* eb = std::getenv("FDFD_BATCH_MEMORY_GB")) budget = std::max(1e-3, std::atof(eb));
    const double usable = std::min(0.4 * (double)freeb, budget * 1073741824.0);
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
        slab_need = (std::max(slab_need, (size_t)B * arena_bytes[ji]) + 255) & ~(size_t)255;
    }
    // two slab halves (factor the next group while solving the current one);
    // one slab: no overlap; none: the factors allocate their arenas one by one
    // (the slab of an earlier batch on this handle is reused when it is large
    // enough, and kept for the next one)
    // FDFD_BATCH_LANES=1 runs the lanes in sequence (one slab half)
    const char* el = std::getenv("FDFD_BATCH_LANES");
    const bool one_lane = el && el[0] == '1';
    int halves = 0;
    for (int want : {2, 1}) {
        if (one_lane && want == 2) continue;
        const size_t bytes = (size_t)want * slab_need;
        if (h->bslab_size >= bytes) {
            halves = want;
            break;
        }
        if (h->bslab) (void)hipFree(h->bslab);
        h->bslab = nullptr;
        h->bslab_size = 0;
        if (hipMalloc((void**)&h->bslab, bytes) == hipSuccess) {
            h->bslab_size = bytes;
            halves = want;
            break;
        }
        (void)hipGetLastError();
        h->bslab = nullptr;
    }
    h->slab = halves ? h->bslab : nullptr;
    // the batch's handle only uses the slab's range (its arena release skips
    // slab pointers); the builder carves the halves
    h->slab_size = (size_t)halves * slab_need;
    h->slab_used = h->slab_size;
    std::vector<Group> groups;
    for (int ji : order)
        for (int o0 = 0; o0 < (int)b->jobs[ji].omegas.size(); o0 += group[ji])
            groups.push_back({ji, o0, std::min(group[ji], (int)b->jobs[ji].omegas.size() - o0)});
    if (timing)
        fprintf(stderr, "batch setup (results %.1f GB, slab %d x %.1f GB, %zu groups): %.3f s\n",
                [&] { double t = 0; for (auto& j : b->jobs) t += (double)j.omegas.size() * j.nsrc * 48.0 * j.g.Nx * j.g.Ny; return t / 1e9; }(),
                halves, slab_need / 1e9, groups.size(), secs(ta0, now()));
    Builder bld;
    if ((rc = builder_init(bld, h))) {
        h->slab = nullptr;
        h->slab_size = h->slab_used = 0;
        free_results(b);
        return rc;
    }
    // eps buffers of both handles and the sources, sized for the largest job
    // up front: a free inside the loop would wait for the other lane's work
    long Nmax = 0, Smax = 0;
    for (const auto& job : b->jobs) {
        Nmax = std::max(Nmax, (long)job.g.Nx * job.g.Ny);
        Smax = std::max(Smax, (long)job.g.Nx * job.g.Ny * job.nsrc);
    }
    double2* dsrc = nullptr;
    for (Handle* H : {h, &bld.H})
        if (rc == FDFD_OK && H->eps_n < Nmax) {
            if (H->eps) (void)hipFree(H->eps);
            H->eps = nullptr;
            H->eps_n = 0;
            if (hipMalloc(&H->eps, sizeof(double2) * Nmax) != hipSuccess)
                rc = fail(h, FDFD_ERR_OOM, "batch: eps buffer allocation failed");
            else
                H->eps_n = Nmax;
        }
    if (rc == FDFD_OK && hipMalloc(&dsrc, sizeof(double2) * Smax) != hipSuccess) {
        dsrc = nullptr;
        rc = fail(h, FDFD_ERR_OOM, "batch: source buffer allocation failed");
    }
    // factor group gi into slot gi % 2 on the builder (any thread)
    std::vector<Factor*> slotF[2];
    double build_s[2] = {0, 0};
    auto build = [&](int gi) -> int {
        const auto tb0 = now();
        (void)hipSetDevice(h->device);
        Handle* H = &bld.H;
        const Group& G = groups[gi];
        BatchJob& job = b->jobs[G.ji];
        const long N = (long)job.g.Nx * job.g.Ny;
        if (bld.job != G.ji) {
            H->g = job.g;
            H->has_grid = true;
            int r = stream_copy(H, H->eps, job.eps.data(), sizeof(double2) * N, hipMemcpyHostToDevice);
            if (r) return r;
            H->has_eps = true;
            bld.job = G.ji;
        }
        const int slot = gi & 1;
        H->slab = halves ? h->slab + (size_t)(halves == 2 ? slot : 0) * slab_need : nullptr;
        H->slab_size = halves ? slab_need : 0;
        H->slab_used = 0;
        std::vector<Factor*>& Fs = slotF[slot];
        Fs.clear();
        for (int q = 0; q < G.nb; ++q) {
            Factor* F = new Factor();
            F->pol = job.pol;
            F->omega = job.omegas[G.o0 + q];
            F->precision = precision;
            Fs.push_back(F);
        }
        const int r = rt_build_factors(H, Fs);
        build_s[slot] = secs(tb0, now());
        return r;
    };
    auto release = [&](int slot) {
        for (Factor* F : slotF[slot]) rt_free_factor(h, F);
        slotF[slot].clear();
    };
    int iters = 0, pivot_fallback = 0;
    double worst = 0, t_factor = 0, t_solve = 0, growth = 0;
    int solve_job = -1;
    std::thread worker;
    int wrc = FDFD_OK;
    int brc = rc == FDFD_OK ? build(0) : FDFD_OK;
    for (int gi = 0; gi < (int)groups.size() && rc == FDFD_OK; ++gi) {
        const int slot = gi & 1;
        const Group& G = groups[gi];
        BatchJob& job = b->jobs[G.ji];
        std::vector<Factor*>& Fs = slotF[slot];
        for (Factor* F : Fs) {
            pivot_fallback |= F->pivot_fallback;
            growth = std::max(growth, F->pivot_growth);
        }
        if (brc) {
            rc = fail(h, brc, bld.H.err);
            break;
        }
        t_factor += Fs[0]->t_factor_ms;
        // the next group factors on the builder while this one is solved
        const bool overlap = halves == 2 && gi + 1 < (int)groups.size();
        if (overlap) worker = std::thread([&, gi] { wrc = build(gi + 1); });
        const auto ts0 = now();
        const GridDesc& g = job.g;
        const long N = (long)g.Nx * g.Ny;
        if (solve_job != G.ji) {
            h->g = g;
            h->has_grid = true;
            rc = stream_copy(h, h->eps, job.eps.data(), sizeof(double2) * N, hipMemcpyHostToDevice);
            h->has_eps = rc == FDFD_OK;
            if (rc == FDFD_OK)
                rc = stream_copy(h, dsrc, job.srcs.data(), sizeof(double2) * N * job.nsrc,
                                 hipMemcpyHostToDevice);
            solve_job = G.ji;
        }
        if (rc == FDFD_OK) {
            hipEvent_t e0, e1;
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, h->stream);
            const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
            for (int q = 0; q < G.nb && rc == FDFD_OK; ++q) {
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
                        double2* out = job.result + ((long)(G.o0 + q) * job.nsrc + s0 + k) * 3 * N;
                        launch_fields(h->stream, out, h->mX + (long)k * N, h->eps, F->sxf, F->sxb,
                                      F->syf, F->syb, g, F->pol, w, 0);
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
        }
        const double solve_s = secs(ts0, now());
        if (overlap) {
            worker.join();
            brc = wrc;
        }
        if (timing)
            fprintf(stderr, "batch group %d (job %d, %d x %d, omegas %d..%d): factor %.3f s%s, solves %.3f s\n",
                    gi, G.ji, g.Nx, g.Ny, G.o0, G.o0 + G.nb - 1, build_s[slot],
                    gi == 0 ? "" : " (beside the previous solves)", solve_s);
        release(slot);  // no build is running: the free synchronizes nothing else
        if (rc) break;
        if (!overlap && gi + 1 < (int)groups.size()) brc = build(gi + 1);
    }
    if (worker.joinable()) worker.join();
    release(0);
    release(1);
    if (dsrc) (void)hipFree(dsrc);
    builder_release_borrowed(bld);
    (void)hipStreamSynchronize(h->stream);
    h->slab = nullptr;  // the slab itself stays with the handle (bslab)
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
