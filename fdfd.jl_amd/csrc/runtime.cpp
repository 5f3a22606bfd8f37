// Host runtime and C ABI of libfdfd: handle lifecycle, operator assembly,
// the twisted block-Thomas factor driver, block sweeps, FP64 flexible GMRES
// refinement, field recovery, adjoint gradient.  Shift-invert Arnoldi is in
// eigen.cpp, batched solves in batch.cpp.
//
// Reference path replaced: solve(d::Device, pol) (src/solver/driven.jl:4-59)
// with dolinearsolve -> lu(A)\b (src/solver/solver.jl:4-41).
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "../../include/fdfd.h"

namespace fdfd {

int fail(Handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail((h), e_ == hipErrorOutOfMemory ? FDFD_ERR_OOM : FDFD_ERR_HIP, \
                        std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

int stream_copy(Handle* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, kind, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FDFD_OK;
}

#define COPY(h, dst, src, bytes, kind)                         \
    do {                                                       \
        int rc_ = stream_copy((h), (dst), (src), (bytes), (kind)); \
        if (rc_) return rc_;                                   \
    } while (0)

// ---------------------------------------------------------------------------
// profiling: HIP events around launches on the handle's stream
// ---------------------------------------------------------------------------
static hipEvent_t ev_get(Prof& p) {
    if (!p.pool.empty()) {
        hipEvent_t e = p.pool.back();
        p.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}
static void prof_resolve(Prof& p) {
    for (auto& q : p.pending) {
        (void)hipEventSynchronize(q.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, q.a, q.b);
        p.ms[q.id] += ms;
        p.launches[q.id] += 1;
        p.flops[q.id] += q.flops;
        p.bytes[q.id] += q.bytes;
        p.pool.push_back(q.a);
        p.pool.push_back(q.b);
    }
    p.pending.clear();
}
struct ProfScope {
    Handle* h; int id; double flops, bytes; hipStream_t st; hipEvent_t a{}, b{};
    ProfScope(Handle* h_, int id_, double f, double by, hipStream_t s = nullptr)
        : h(h_), id(id_), flops(f), bytes(by), st(s ? s : h_->stream) {
        if (h->prof.enabled) {
            a = ev_get(h->prof);
            (void)hipEventRecord(a, st);
        }
    }
    ~ProfScope() {
        if (h->prof.enabled) {
            b = ev_get(h->prof);
            (void)hipEventRecord(b, st);
            h->prof.pending.push_back({a, b, id, flops, bytes});
            if (h->prof.pending.size() > 4096) prof_resolve(h->prof);
        }
    }
};

// ---------------------------------------------------------------------------
// work buffers
// ---------------------------------------------------------------------------
static void dfree(void* p) {
    if (p) (void)hipFree(p);
}

static void clear_graphs(Handle* h) {
    for (auto& g : h->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        dfree(g.lj_d);
        dfree(g.sj_d);
    }
    h->graphs.clear();
}

static int ensure_factor_work(Handle* h, int n, int nch) {
    if (h->work_n >= n && h->work_ch >= nch) return FDFD_OK;
    n = std::max(n, h->work_n);
    nch = std::max(nch, h->work_ch);
    clear_graphs(h);
    dfree(h->used); dfree(h->q); dfree(h->piv); dfree(h->Vbuf); dfree(h->LU); dfree(h->Zb);
    dfree(h->win); dfree(h->verr);
    h->used = h->q = h->piv = h->win = h->verr = nullptr;
    h->Vbuf = h->LU = h->Zb = nullptr;
    const int NBmax = PMAX;
    HIPCHK(h, hipMalloc(&h->used, sizeof(int) * nch * n));
    HIPCHK(h, hipMalloc(&h->q, sizeof(int) * nch * n));
    HIPCHK(h, hipMalloc(&h->piv, sizeof(int) * nch * NBmax));
    HIPCHK(h, hipMalloc(&h->win, sizeof(int) * nch * WIN));
    HIPCHK(h, hipMalloc(&h->verr, sizeof(int)));
    HIPCHK(h, hipMalloc(&h->Vbuf, sizeof(double2) * nch * (size_t)n * NBmax));
    HIPCHK(h, hipMalloc(&h->LU, sizeof(double2) * nch * 2 * NBmax * NBmax));
    HIPCHK(h, hipMalloc(&h->Zb, sizeof(double2) * nch * (size_t)NBmax * n));
    h->work_n = n;
    h->work_ch = nch;
    return FDFD_OK;
}

static const int KDIM = 30;

static int ensure_solve_work(Handle* h, long N) {
    if (h->work_N >= N) return FDFD_OK;
    void** bufs[] = {(void**)&h->r_in, (void**)&h->g_mid, (void**)&h->x_out, (void**)&h->V,
                     (void**)&h->Zk, (void**)&h->tmp, (void**)&h->tmp2, (void**)&h->bvec,
                     (void**)&h->xvec};
    for (auto b : bufs) { dfree(*b); *b = nullptr; }
    HIPCHK(h, hipMalloc(&h->r_in, sizeof(double2) * N));
    HIPCHK(h, hipMalloc(&h->g_mid, sizeof(double2) * N));
    HIPCHK(h, hipMalloc(&h->x_out, sizeof(double2) * N));
    HIPCHK(h, hipMalloc(&h->V, sizeof(double2) * N * (KDIM + 1)));
    HIPCHK(h, hipMalloc(&h->Zk, sizeof(double2) * N * KDIM));
    HIPCHK(h, hipMalloc(&h->tmp, sizeof(double2) * N));
    HIPCHK(h, hipMalloc(&h->tmp2, sizeof(double2) * N));
    HIPCHK(h, hipMalloc(&h->bvec, sizeof(double2) * N));
    HIPCHK(h, hipMalloc(&h->xvec, sizeof(double2) * N));
    if (!h->dpart) {
        h->nblk = 512;
        HIPCHK(h, hipMalloc(&h->dpart, sizeof(double2) * h->nblk * (KDIM + 2)));
        HIPCHK(h, hipMalloc(&h->rpart, sizeof(double) * h->nblk));
        HIPCHK(h, hipMalloc(&h->hdev, sizeof(double2) * (KDIM + 2)));
        HIPCHK(h, hipMalloc(&h->tpart, sizeof(double2) * 2 * 16 * MAX_LINE));
    }
    h->work_N = N;
    h->kdim = KDIM;
    return FDFD_OK;
}

static void free_factor(Factor* F) {
    if (!F) return;
    dfree(F->coef); dfree(F->sxf); dfree(F->sxb); dfree(F->syf); dfree(F->syb);
    dfree(F->Y); dfree(F->rho); dfree(F->rhoinv); dfree(F->kap); dfree(F->kapinv);
    dfree(F->pivstat); dfree(F->sweep_tab[0]); dfree(F->sweep_tab[1]);
    for (auto& e : F->sweep_exec)
        if (e) (void)hipGraphExecDestroy(e);
    delete F;
}

// ---------------------------------------------------------------------------
// operator assembly
// ---------------------------------------------------------------------------
static int assemble(Handle* h, Factor* F) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
    const double eta0 = std::sqrt(MU0 / EPS0);
    HIPCHK(h, hipMalloc(&F->coef, sizeof(double2) * 5 * N));
    HIPCHK(h, hipMalloc(&F->sxf, sizeof(double2) * g.Nx));
    HIPCHK(h, hipMalloc(&F->sxb, sizeof(double2) * g.Nx));
    HIPCHK(h, hipMalloc(&F->syf, sizeof(double2) * g.Ny));
    HIPCHK(h, hipMalloc(&F->syb, sizeof(double2) * g.Ny));
    launch_sfactor(h->stream, F->sxf, F->sxb, F->syf, F->syb, g, F->omega, eps0n, eta0);
    double kdiag;
    if (F->pol == FDFD_TM) {
        F->scale = mu0n * g.dx * g.dx;
        kdiag = F->omega * F->omega * eps0n * mu0n * g.dx * g.dx;
    } else {
        F->scale = eps0n * g.dx * g.dx;
        kdiag = F->omega * F->omega * mu0n * eps0n * g.dx * g.dx;
    }
    {
        ProfScope ps(h, P_STENCIL, 0, 0);
        launch_coeffs(h->stream, F->coef, h->eps, F->sxf, F->sxb, F->syf, F->syb, g, F->pol,
                      kdiag);
    }
    HIPCHK(h, hipGetLastError());
    return FDFD_OK;
}

// coefficient plane pointers for the line orientation
struct Planes { const double2 *L, *U, *sub, *sup, *diag; };
static Planes planes(const Factor* F, long N) {
    const double2* c = F->coef;
    if (F->orient == 0) return {c + CS * N, c + CN * N, c + CW * N, c + CE * N, c + CC * N};
    return {c + CW * N, c + CE * N, c + CS * N, c + CN * N, c + CC * N};
}

static void* yline(const Factor* F, int j) {
    return (char*)F->Y + (size_t)j * F->n * F->n * F->elem;
}

// ---------------------------------------------------------------------------
// sweep tables (see the twisted block-Thomas recurrences in build_factor)
// ---------------------------------------------------------------------------
static int build_sweep_tables(Handle* h, Factor* F) {
    const long N = (long)h->g.Nx * h->g.Ny;
    const int m = F->m, n = F->n, k = F->k;
    const int nsteps = std::max(k, m - 1 - k);
    Planes P = planes(F, N);
    const long sj = F->sj;
    auto base = [&](const int j, const int trans) {
        SweepJob J{};
        J.Y = yline(F, j);
        J.rho = F->rho + (long)j * n; J.rhoinv = F->rhoinv + (long)j * n;
        J.kap = F->kap + (long)j * n; J.kapinv = F->kapinv + (long)j * n;
        J.si = F->si; J.n = n;
        (void)trans;
        return J;
    };
    F->nfwd = nsteps;
    F->nbwd = nsteps;
    F->fwd_jobs.assign(nsteps, 0);
    F->bwd_jobs.assign(nsteps, 0);
    F->fwd_has_a.assign(nsteps, 0);
    F->bwd_has_a.assign(nsteps, 0);
    for (int tr = 0; tr < 2; ++tr) {
        std::vector<SweepJob> tab((size_t)2 * (2 * nsteps + 1));
        for (int s = 0; s < nsteps; ++s) {
            int nj = 0;
            if (s < k) {  // forward A, line j = s
                int j = s;
                SweepJob J = base(j, tr);
                J.mode = 0;
                J.b = h->r_in + j * sj;
                if (j > 0) {
                    J.nb1 = h->g_mid + (j - 1) * sj;
                    J.c1 = tr ? P.U + (j - 1) * sj : P.L + j * sj;
                }
                J.out = h->g_mid + j * sj;
                tab[2 * s + nj++] = J;
                F->fwd_has_a[s] = 1;
            }
            if (s < m - 1 - k) {  // forward B, line j = m-1-s
                int j = m - 1 - s;
                SweepJob J = base(j, tr);
                J.mode = 0;
                J.b = h->r_in + j * sj;
                if (j < m - 1) {
                    J.nb1 = h->g_mid + (j + 1) * sj;
                    J.c1 = tr ? P.L + (j + 1) * sj : P.U + j * sj;
                }
                J.out = h->g_mid + j * sj;
                tab[2 * s + nj++] = J;
            }
            F->fwd_jobs[s] = nj;
        }
        {  // middle line k
            SweepJob J = base(k, tr);
            J.mode = 2;
            J.b = h->r_in + k * sj;
            if (k >= 1) {
                J.nb1 = h->g_mid + (k - 1) * sj;
                J.c1 = tr ? P.U + (k - 1) * sj : P.L + k * sj;
            }
            if (k + 1 <= m - 1) {
                J.nb2 = h->g_mid + (k + 1) * sj;
                J.c2 = tr ? P.L + (k + 1) * sj : P.U + k * sj;
            }
            J.out = h->x_out + k * sj;
            tab[2 * nsteps] = J;
        }
        for (int s = 0; s < nsteps; ++s) {
            int nj = 0;
            int ja = k - 1 - s;
            if (ja >= 0) {  // backward A
                SweepJob J = base(ja, tr);
                J.mode = 1;
                J.nb1 = h->x_out + (ja + 1) * sj;
                J.c1 = tr ? P.L + (ja + 1) * sj : P.U + ja * sj;
                J.g = h->g_mid + ja * sj;
                J.out = h->x_out + ja * sj;
                tab[2 * (nsteps + 1 + s) + nj++] = J;
                F->bwd_has_a[s] = 1;
            }
            int jb = k + 1 + s;
            if (jb <= m - 1) {  // backward B
                SweepJob J = base(jb, tr);
                J.mode = 1;
                J.nb1 = h->x_out + (jb - 1) * sj;
                J.c1 = tr ? P.U + (jb - 1) * sj : P.L + jb * sj;
                J.g = h->g_mid + jb * sj;
                J.out = h->x_out + jb * sj;
                tab[2 * (nsteps + 1 + s) + nj++] = J;
            }
            F->bwd_jobs[s] = nj;
        }
        HIPCHK(h, hipMalloc(&F->sweep_tab[tr], sizeof(SweepJob) * tab.size()));
        COPY(h, F->sweep_tab[tr], tab.data(), sizeof(SweepJob) * tab.size(),
             hipMemcpyHostToDevice);
    }
    return FDFD_OK;
}

// out = M^-1 in (transpose: M^-T in), M = block-tridiagonal part of A' without
// the periodic corner blocks between the first and last line.
static void launch_sweeps(Handle* h, Factor* F, int trans, bool prof) {
    const int n = F->n;
    const SweepJob* tab = F->sweep_tab[trans];
    const int nsteps = F->nfwd;
    const double bpl = (double)n * n * F->elem;  // factor bytes per line
    auto run = [&](hipStream_t st, const SweepJob* t, double2* part) {
        ProfScope* ps = prof ? new ProfScope(h, P_SWEEP, 8.0 * n * n, bpl, st) : nullptr;
        if (F->precision == FDFD_FP64) launch_sweep<double>(st, t, 1, n, trans, part);
        else launch_sweep<float>(st, t, 1, n, trans, part);
        delete ps;
    };
    // transposed sweeps keep their partial sums in one buffer slot per stream
    double2* partB = h->tpart + (size_t)16 * MAX_LINE;
    auto fork = [&]() {
        (void)hipEventRecord(h->ev_fork, h->stream);
        (void)hipStreamWaitEvent(h->stream2, h->ev_fork, 0);
    };
    auto join = [&]() {
        (void)hipEventRecord(h->ev_join, h->stream2);
        (void)hipStreamWaitEvent(h->stream, h->ev_join, 0);
    };
    // chain A (listed first in a step) on the handle's stream, chain B on the
    // second stream; the middle line needs both forward sweeps
    fork();
    for (int s = 0; s < nsteps; ++s) {
        const SweepJob* t = tab + 2 * s;
        if (F->fwd_has_a[s]) run(h->stream, t, h->tpart);
        if (F->fwd_jobs[s] > F->fwd_has_a[s]) run(h->stream2, t + F->fwd_has_a[s], partB);
    }
    join();
    run(h->stream, tab + 2 * nsteps, h->tpart);
    fork();
    for (int s = 0; s < nsteps; ++s) {
        const SweepJob* t = tab + 2 * (nsteps + 1 + s);
        if (F->bwd_has_a[s]) run(h->stream, t, h->tpart);
        if (F->bwd_jobs[s] > F->bwd_has_a[s]) run(h->stream2, t + F->bwd_has_a[s], partB);
    }
    join();
}

// out = M^-1 in (transpose: M^-T in), M = block-tridiagonal part of A' without
// the periodic corner blocks between the first and last line.  The sweep
// launch sequence of a factor is captured once in a HIP graph and replayed
// (direct launches while per-kernel profiling is on).
static int prec_apply(Handle* h, Factor* F, int trans, const double2* in, double2* out) {
    const long N = (long)h->g.Nx * h->g.Ny;
    HIPCHK(h, hipMemcpyAsync(h->r_in, in, sizeof(double2) * N, hipMemcpyDeviceToDevice,
                             h->stream));
    if (h->prof.enabled || !h->use_graphs) {
        launch_sweeps(h, F, trans, h->prof.enabled);
    } else {
        if (!F->sweep_exec[trans]) {
            hipGraph_t gr;
            HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            launch_sweeps(h, F, trans, false);
            HIPCHK(h, hipStreamEndCapture(h->stream, &gr));
            hipError_t e = hipGraphInstantiate(&F->sweep_exec[trans], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                F->sweep_exec[trans] = nullptr;
                return fail(h, FDFD_ERR_HIP, std::string("sweep graph: ") + hipGetErrorString(e));
            }
        }
        HIPCHK(h, hipGraphLaunch(F->sweep_exec[trans], h->stream));
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(out, h->x_out, sizeof(double2) * N, hipMemcpyDeviceToDevice,
                             h->stream));
    return FDFD_OK;
}

// ---------------------------------------------------------------------------
// twisted block-Thomas factorization
//   chain A (lines 0..k-1, upward):   S_j = D_j - L_j X_{j-1} U_{j-1}
//   chain B (lines m-1..k+1, down):   S_j = D_j - U_j X_{j+1} L_{j+1}
//   middle line k:                    S_k = D_k - L_k X_{k-1} U_{k-1} - U_k X_{k+1} L_{k+1}
// X_j = S_j^-1 by in-place Gauss-Jordan with partial pivoting.
// ---------------------------------------------------------------------------
static int build_factors(Handle* h, const std::vector<Factor*>& Fs) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const int B = (int)Fs.size();
    if (B < 1) return fail(h, FDFD_ERR_INVALID, "no factor to build");
    const int prec = Fs[0]->precision;
    for (Factor* F : Fs) {
        if (F->precision != prec)
            return fail(h, FDFD_ERR_INVALID, "factors built together share one precision");
        if (g.Nx <= g.Ny) { F->orient = 0; F->n = g.Nx; F->m = g.Ny; F->si = 1; F->sj = g.Nx; }
        else { F->orient = 1; F->n = g.Ny; F->m = g.Nx; F->si = g.Nx; F->sj = 1; }
        F->elem = (F->precision == FDFD_FP64) ? sizeof(double2) : sizeof(float2);
        F->k = F->m / 2;
    }
    const Factor* F0 = Fs[0];
    const int n = F0->n, m = F0->m, k = F0->k;
    const size_t elem = F0->elem;
    if (n > MAX_LINE)
        return fail(h, FDFD_ERR_INVALID, "line block larger than 1024 cells (min(Nx,Ny) > 1024)");
    int rc;
    if ((rc = ensure_factor_work(h, n, 2 * B))) return rc;
    if ((rc = ensure_solve_work(h, N))) return rc;
    for (Factor* F : Fs) {
        if ((rc = assemble(h, F))) return rc;
        HIPCHK(h, hipMalloc(&F->Y, (size_t)m * n * n * elem));
        HIPCHK(h, hipMalloc(&F->rho, sizeof(int) * (size_t)m * n));
        HIPCHK(h, hipMalloc(&F->rhoinv, sizeof(int) * (size_t)m * n));
        HIPCHK(h, hipMalloc(&F->kap, sizeof(int) * (size_t)m * n));
        HIPCHK(h, hipMalloc(&F->kapinv, sizeof(int) * (size_t)m * n));
        HIPCHK(h, hipMalloc(&F->pivstat, sizeof(double) * 3 * m));
    }
    std::vector<double> ps0(3 * (size_t)m);
    for (int j = 0; j < m; ++j) {
        ps0[3 * j] = std::numeric_limits<double>::infinity();
        ps0[3 * j + 1] = 0;
        ps0[3 * j + 2] = 0;
    }
    const int NB = h->panel_width ? h->panel_width : PMAX;  // fast-panel width
    const int nsteps = std::max(k, m - 1 - k);
    const int JS = 2 * B;  // job slots per step
    const long sj = F0->sj;
    std::vector<LineJob> lj((size_t)JS * (nsteps + 1));
    std::vector<SchurJob> sjb((size_t)JS * (nsteps + 1));
    std::vector<int> njobs(nsteps + 1, 0);
    // chain c (2 per factor) owns slices of the handle's work buffers
    auto line_job = [&](const Factor* F, int j, int c) {
        LineJob J{};
        J.Y = yline(F, j);
        J.used = h->used + (long)c * n;
        J.prow = F->rho + (long)j * n;
        J.iprow = F->rhoinv + (long)j * n;
        J.q = h->q + (long)c * n;
        J.kap = F->kap + (long)j * n;
        J.kapinv = F->kapinv + (long)j * n;
        J.Vbuf = (char*)h->Vbuf + (size_t)c * n * PMAX * elem;
        J.LU = (char*)h->LU + (size_t)c * 2 * PMAX * PMAX * elem;
        J.piv = h->piv + c * PMAX;
        J.Z = (char*)h->Zb + (size_t)c * PMAX * n * elem;
        J.win = h->win + c * WIN;
        J.verr = h->verr;
        J.pivstat = F->pivstat + 3 * (long)j;
        J.n = n;
        return J;
    };
    auto schur_base = [&](const Factor* F, const Planes& P, int j, int c) {
        SchurJob S{};
        S.S = yline(F, j);
        S.dsub = P.sub + j * sj; S.ddiag = P.diag + j * sj; S.dsup = P.sup + j * sj;
        S.si = F->si;
        S.q = h->q + (long)c * n;
        S.used = h->used + (long)c * n;
        S.iprow = F->rhoinv + (long)j * n;
        S.win = h->win + c * WIN;
        S.n = n;
        return S;
    };
    // per step: the chain-A jobs of all factors, then the chain-B jobs
    std::vector<int> nA(nsteps + 1, 0), nB(nsteps + 1, 0);
    for (int s = 0; s <= nsteps; ++s) {
        int nj = 0;
        for (int chain = 0; chain < 2; ++chain) {
            for (int f = 0; f < B; ++f) {
                const Factor* F = Fs[f];
                const Planes P = planes(F, N);
                if (s == nsteps) {  // middle line k (listed once)
                    if (chain == 1) continue;
                    const int c = 2 * f;
                    lj[(size_t)JS * s + nj] = line_job(F, k, c);
                    SchurJob S = schur_base(F, P, k, c);
                    if (k >= 1) {
                        S.Y1 = yline(F, k - 1);
                        S.rhoinv1 = F->rhoinv + (long)(k - 1) * n;
                        S.kap1 = F->kap + (long)(k - 1) * n;
                        S.l1 = P.L + k * sj;
                        S.u1 = P.U + (k - 1) * sj;
                    }
                    if (k + 1 <= m - 1) {
                        S.Y2 = yline(F, k + 1);
                        S.rho2 = F->rho + (long)(k + 1) * n;
                        S.kap2 = F->kap + (long)(k + 1) * n;
                        S.l2 = P.U + k * sj;
                        S.u2 = P.L + (k + 1) * sj;
                    }
                    sjb[(size_t)JS * s + nj] = S;
                    ++nj;
                    ++nA[s];
                    continue;
                }
                if (chain == 0 && s < k) {  // chain A, line j = s
                    const int j = s, c = 2 * f;
                    lj[(size_t)JS * s + nj] = line_job(F, j, c);
                    SchurJob S = schur_base(F, P, j, c);
                    if (j > 0) {  // S_j = D_j - L_j X_{j-1} U_{j-1}
                        S.Y1 = yline(F, j - 1);
                        S.rhoinv1 = F->rhoinv + (long)(j - 1) * n;
                        S.kap1 = F->kap + (long)(j - 1) * n;
                        S.l1 = P.L + j * sj;
                        S.u1 = P.U + (j - 1) * sj;
                    }
                    sjb[(size_t)JS * s + nj] = S;
                    ++nj;
                    ++nA[s];
                }
                if (chain == 1 && s < m - 1 - k) {  // chain B, line j = m-1-s
                    const int j = m - 1 - s, c = 2 * f + 1;
                    lj[(size_t)JS * s + nj] = line_job(F, j, c);
                    SchurJob S = schur_base(F, P, j, c);
                    if (j < m - 1) {  // S_j = D_j - U_j X_{j+1} L_{j+1}
                        S.Y1 = yline(F, j + 1);
                        S.rhoinv1 = F->rhoinv + (long)(j + 1) * n;
                        S.kap1 = F->kap + (long)(j + 1) * n;
                        S.l1 = P.U + j * sj;
                        S.u1 = P.L + (j + 1) * sj;
                    }
                    sjb[(size_t)JS * s + nj] = S;
                    ++nj;
                    ++nB[s];
                }
            }
        }
        njobs[s] = nj;
    }
    const bool fp64 = (prec == FDFD_FP64);
    // one pass of the factorization as a launch sequence
    // chain A runs on the handle's stream and chain B on the second stream
    // (they are independent until the middle line); joined before the middle
    auto launch_pass = [&](const LineJob* lj_d, const SchurJob* sj_d, bool full, bool prof) {
        auto line = [&](hipStream_t st, const LineJob* J, const SchurJob* SJ, int nj) {
            auto scope = [&](int id, double fl, double by) {
                return prof ? new ProfScope(h, id, fl, by, st) : nullptr;
            };
            {
                ProfScope* ps = scope(P_SCHUR, 0, 2.0 * n * n * elem * nj);
                if (fp64) launch_schur<double>(st, SJ, nj, n);
                else launch_schur<float>(st, SJ, nj, n);
                delete ps;
            }
            for (int k0 = 0; k0 < n;) {
                int nb;
                const int fresh = std::min(n, WIN + k0);
                {
                    ProfScope* ps = scope(P_PANEL, 0, 0);
                    if (!full) {
                        nb = std::min(NB, n - k0);
                        if (fp64) launch_wpanel<double>(st, J, nj, n, k0, nb, fresh);
                        else launch_wpanel<float>(st, J, nj, n, k0, nb, fresh);
                    } else {
                        nb = 16;  // full-search panel width: powers of two
                        while (nb > n - k0) nb >>= 1;
                        if (fp64) launch_panel<double>(st, J, nj, n, k0, nb);
                        else launch_panel<float>(st, J, nj, n, k0, nb);
                    }
                    delete ps;
                }
                {
                    ProfScope* ps = scope(P_ZBUILD, 0, 0);
                    if (fp64) launch_zbuild<double>(st, J, nj, n, k0, nb, fresh, !full);
                    else launch_zbuild<float>(st, J, nj, n, k0, nb, fresh, !full);
                    delete ps;
                }
                {
                    ProfScope* ps = scope(P_TRAIL, 8.0 * n * (double)n * nb * nj,
                                          2.0 * n * (double)n * elem * nj);
                    if (fp64) launch_trailing<double>(st, J, nj, n, k0, nb);
                    else launch_trailing<float>(st, J, nj, n, k0, nb);
                    delete ps;
                }
                k0 += nb;
            }
            launch_maps(st, J, nj, n);
        };
        (void)hipEventRecord(h->ev_fork, h->stream);
        (void)hipStreamWaitEvent(h->stream2, h->ev_fork, 0);
        for (int s = 0; s < nsteps; ++s) {
            const LineJob* J = lj_d + (size_t)JS * s;
            const SchurJob* SJ = sj_d + (size_t)JS * s;
            if (nA[s]) line(h->stream, J, SJ, nA[s]);
            if (nB[s]) line(h->stream2, J + nA[s], SJ + nA[s], nB[s]);
        }
        (void)hipEventRecord(h->ev_join, h->stream2);
        (void)hipStreamWaitEvent(h->stream, h->ev_join, 0);
        line(h->stream, lj_d + (size_t)JS * nsteps, sj_d + (size_t)JS * nsteps, njobs[nsteps]);
    };

    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    // Pass 0: windowed pivot search with verification.  Pass 1 (only if a
    // row outside a window would have won a GEPP step): full-column search.
    int used_full = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool full = (pass == 1) || h->full_pivot_search;
        used_full = full ? 1 : 0;
        // job tables live at fixed addresses per shape so that a captured
        // launch sequence can be replayed for every factorization of it
        FactorGraph* G = nullptr;
        for (auto& gr : h->graphs)
            if (gr.n == n && gr.m == m && gr.prec == prec && gr.nb == NB && gr.full == full &&
                gr.nfac == B)
                G = &gr;
        if (!G) {
            h->graphs.push_back(FactorGraph{});
            G = &h->graphs.back();
            G->n = n; G->m = m; G->prec = prec; G->nb = NB; G->full = full; G->nfac = B;
            HIPCHK(h, hipMalloc(&G->lj_d, sizeof(LineJob) * lj.size()));
            HIPCHK(h, hipMalloc(&G->sj_d, sizeof(SchurJob) * sjb.size()));
        }
        HIPCHK(h, hipMemcpyAsync(G->lj_d, lj.data(), sizeof(LineJob) * lj.size(),
                                 hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(G->sj_d, sjb.data(), sizeof(SchurJob) * sjb.size(),
                                 hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemsetAsync(h->verr, 0, sizeof(int), h->stream));
        for (Factor* F : Fs)
            HIPCHK(h, hipMemcpyAsync(F->pivstat, ps0.data(), sizeof(double) * ps0.size(),
                                     hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));  // host tables are stack-local
        HIPCHK(h, hipEventRecord(e0, h->stream));
        if (h->prof.enabled || !h->use_graphs) {
            launch_pass(G->lj_d, G->sj_d, full, h->prof.enabled);
        } else {
            if (!G->exec) {
                hipGraph_t gr;
                HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                launch_pass(G->lj_d, G->sj_d, full, false);
                HIPCHK(h, hipStreamEndCapture(h->stream, &gr));
                hipError_t e = hipGraphInstantiate(&G->exec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) {
                    G->exec = nullptr;
                    return fail(h, FDFD_ERR_HIP, std::string("factor graph: ") + hipGetErrorString(e));
                }
            }
            HIPCHK(h, hipGraphLaunch(G->exec, h->stream));
        }
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventRecord(e1, h->stream));
        HIPCHK(h, hipEventSynchronize(e1));
        int verr = 0;
        COPY(h, &verr, h->verr, sizeof(int), hipMemcpyDeviceToHost);
        if (full || !verr) break;
    }
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int first_singular = -1;
    for (Factor* F : Fs) {
        F->t_factor_ms = ms;
        F->pivot_fallback = used_full;
        std::vector<double> ps(3 * (size_t)m);
        COPY(h, ps.data(), F->pivstat, sizeof(double) * ps.size(), hipMemcpyDeviceToHost);
        double mn = std::numeric_limits<double>::infinity(), mx = 0;
        F->singular_line = -1;
        for (int j = 0; j < m; ++j) {
            mn = std::min(mn, ps[3 * j]);
            mx = std::max(mx, ps[3 * j + 1]);
            if (ps[3 * j + 2] != 0 && F->singular_line < 0) F->singular_line = j;
        }
        F->pivot_growth = (mn > 0) ? mx / mn : std::numeric_limits<double>::infinity();
        if (F->singular_line >= 0 && first_singular < 0) first_singular = F->singular_line;
    }
    if (first_singular >= 0)
        return fail(h, FDFD_ERR_SINGULAR, "zero pivot in line block " + std::to_string(first_singular));
    for (Factor* F : Fs)
        if ((rc = build_sweep_tables(h, F))) return rc;
    return FDFD_OK;
}

static int build_factor(Handle* h, Factor* F) { return build_factors(h, std::vector<Factor*>{F}); }

// ---------------------------------------------------------------------------
// FP64 Krylov building blocks
// ---------------------------------------------------------------------------
static int dev_norm(Handle* h, const double2* x, long N, double* out) {
    launch_norm2(h->stream, x, N, h->rpart, h->nblk);
    launch_reduce_real(h->stream, h->rpart, h->nblk, (double*)h->hdev);
    double v = 0;
    HIPCHK(h, hipMemcpyAsync(&v, h->hdev, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *out = std::sqrt(v);
    return FDFD_OK;
}

// h[0..k-1] = V[0..k-1]^H w
// (in chunks of at most 32 vectors, the multidot kernel's register budget)
static int dev_multidot(Handle* h, const double2* V, long N, int k, const double2* w, cd* out) {
    for (int c0 = 0; c0 < k; c0 += 32) {
        const int kc = std::min(32, k - c0);
        launch_multidot(h->stream, V + (long)c0 * N, N, kc, w, N, h->dpart, h->nblk);
        launch_reduce_partials(h->stream, h->dpart, h->nblk, kc, h->hdev);
        HIPCHK(h, hipMemcpyAsync(out + c0, h->hdev, sizeof(double2) * kc, hipMemcpyDeviceToHost,
                                 h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return FDFD_OK;
}

// w += sign * V[0..k-1] c   (in chunks of at most 32 vectors)
static int dev_update(Handle* h, double2* w, const double2* V, long N, int k, const cd* c,
                      double sign) {
    for (int c0 = 0; c0 < k; c0 += 32) {
        const int kc = std::min(32, k - c0);
        HIPCHK(h, hipMemcpyAsync(h->hdev, c + c0, sizeof(double2) * kc, hipMemcpyHostToDevice,
                                 h->stream));
        launch_gemv_update(h->stream, w, V + (long)c0 * N, N, kc, h->hdev, sign, N);
        HIPCHK(h, hipStreamSynchronize(h->stream));  // c lives in host memory
    }
    return FDFD_OK;
}

struct SolveStats { int iters = 0, restarts = 0; double resid = 0; };

// Right-preconditioned flexible GMRES(KDIM) on A' x = b (or A'^T x = b) in
// FP64, preconditioned by the block factors.  x is the initial guess on entry.
static int gmres(Handle* h, Factor* F, int trans, const double2* b, double2* x, double tol,
                 int max_restarts, SolveStats& st) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const int K = h->kdim;
    double2* V = h->V;
    double2* Zk = h->Zk;
    double bnorm;
    int rc;
    if ((rc = dev_norm(h, b, N, &bnorm))) return rc;
    if (bnorm == 0) {
        launch_zero(h->stream, x, N);
        st.resid = 0;
        return FDFD_OK;
    }
    std::vector<cd> H((size_t)(K + 1) * K), gv(K + 1), sn(K), hc(K + 1), hc2(K + 1), y(K);
    std::vector<double> cs(K);
    double resid = 1;
    for (int rs = 0; rs <= max_restarts; ++rs) {
        {
            ProfScope ps(h, P_STENCIL, 0, 64.0 * N);
            launch_apply(h->stream, V, x, F->coef, g.Nx, g.Ny, trans, b);
        }
        double beta;
        if ((rc = dev_norm(h, V, N, &beta))) return rc;
        resid = beta / bnorm;
        st.resid = resid;
        if (resid <= tol) return FDFD_OK;
        if (rs == max_restarts) break;
        if (rs > 0) st.restarts++;
        launch_scale_copy(h->stream, V, V, make_double2(1.0 / beta, 0), N);
        std::fill(gv.begin(), gv.end(), cd(0));
        gv[0] = beta;
        int jdone = 0;
        for (int j = 0; j < K; ++j) {
            double2* vj = V + (long)j * N;
            double2* zj = Zk + (long)j * N;
            double2* w = V + (long)(j + 1) * N;
            if ((rc = prec_apply(h, F, trans, vj, zj))) return rc;
            {
                ProfScope ps(h, P_STENCIL, 0, 64.0 * N);
                launch_apply(h->stream, w, zj, F->coef, g.Nx, g.Ny, trans, nullptr);
            }
            // CGS2
            if ((rc = dev_multidot(h, V, N, j + 1, w, hc.data()))) return rc;
            if ((rc = dev_update(h, w, V, N, j + 1, hc.data(), -1.0))) return rc;
            if ((rc = dev_multidot(h, V, N, j + 1, w, hc2.data()))) return rc;
            if ((rc = dev_update(h, w, V, N, j + 1, hc2.data(), -1.0))) return rc;
            for (int i = 0; i <= j; ++i) H[(size_t)i * K + j] = hc[i] + hc2[i];
            double hn;
            if ((rc = dev_norm(h, w, N, &hn))) return rc;
            H[(size_t)(j + 1) * K + j] = hn;
            if (hn > 0) launch_scale_copy(h->stream, w, w, make_double2(1.0 / hn, 0), N);
            // Givens
            for (int i = 0; i < j; ++i) {
                cd a = H[(size_t)i * K + j], bb = H[(size_t)(i + 1) * K + j];
                H[(size_t)i * K + j] = cs[i] * a + sn[i] * bb;
                H[(size_t)(i + 1) * K + j] = -std::conj(sn[i]) * a + cs[i] * bb;
            }
            cd a = H[(size_t)j * K + j];
            double bb = hn;
            double t = std::sqrt(std::norm(a) + bb * bb);
            if (std::abs(a) == 0) {
                cs[j] = 0;
                sn[j] = 1;
                H[(size_t)j * K + j] = bb;
            } else {
                cd ph = a / std::abs(a);
                cs[j] = std::abs(a) / t;
                sn[j] = ph * bb / t;
                H[(size_t)j * K + j] = ph * t;
            }
            gv[j + 1] = -std::conj(sn[j]) * gv[j];
            gv[j] = cs[j] * gv[j];
            st.iters++;
            jdone = j + 1;
            double est = std::abs(gv[j + 1]) / bnorm;
            if (est <= 0.5 * tol || hn == 0) break;
        }
        // y = H^-1 g, x += Z y
        for (int i = jdone - 1; i >= 0; --i) {
            cd s = gv[i];
            for (int l = i + 1; l < jdone; ++l) s -= H[(size_t)i * K + l] * y[l];
            y[i] = s / H[(size_t)i * K + i];
        }
        if ((rc = dev_update(h, x, Zk, N, jdone, y.data(), 1.0))) return rc;
    }
    st.resid = resid;
    char buf[160];
    snprintf(buf, sizeof(buf),
             "refinement did not reach the tolerance %.1e (relative residual %.3e)", tol, resid);
    return fail(h, FDFD_ERR_NOCONV, buf);
}

}  // namespace fdfd

using namespace fdfd;

// ===========================================================================
// C ABI
// ===========================================================================
static Handle* H(fdfd_handle* h) { return h ? &h->impl : nullptr; }

static Factor* find_factor(Handle* h, int id) {
    auto it = h->factors.find(id);
    return it == h->factors.end() ? nullptr : it->second;
}

extern "C" {

int fdfd_version(int* major, int* minor, int* patch) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    if (patch) *patch = 0;
    return FDFD_OK;
}

const char* fdfd_last_error(const fdfd_handle* h) { return h ? h->impl.err.c_str() : "null handle"; }

const char* fdfd_arch(void) { return "gfx950"; }

int fdfd_create(int device_ordinal, fdfd_handle** out) {
    if (!out) return FDFD_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return FDFD_ERR_HIP;
    if (device_ordinal < 0 || device_ordinal >= ndev) return FDFD_ERR_INVALID;
    if (hipSetDevice(device_ordinal) != hipSuccess) return FDFD_ERR_HIP;
    auto* hh = new fdfd_handle();
    hh->impl.device = device_ordinal;
    const char* eg = std::getenv("FDFD_GRAPHS");
    hh->impl.use_graphs = !(eg && eg[0] == '0');
    if (hipStreamCreateWithFlags(&hh->impl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&hh->impl.stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&hh->impl.ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&hh->impl.ev_join, hipEventDisableTiming) != hipSuccess) {
        delete hh;
        return FDFD_ERR_HIP;
    }
    *out = hh;
    return FDFD_OK;
}

int fdfd_destroy(fdfd_handle* hh) {
    Handle* h = H(hh);
    if (!h) return FDFD_ERR_INVALID;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    for (auto& kv : h->factors) free_factor(kv.second);
    clear_graphs(h);
    void* bufs[] = {h->eps, h->used, h->q, h->piv, h->win, h->verr, h->Vbuf, h->LU, h->Zb,
                    h->r_in, h->g_mid,
                    h->x_out, h->V, h->Zk, h->tmp, h->tmp2, h->bvec, h->xvec, h->dpart,
                    h->rpart, h->hdev, h->tpart};
    for (void* b : bufs) dfree(b);
    prof_resolve(h->prof);
    for (auto e : h->prof.pool) (void)hipEventDestroy(e);
    (void)hipEventDestroy(h->ev_fork);
    (void)hipEventDestroy(h->ev_join);
    (void)hipStreamDestroy(h->stream2);
    (void)hipStreamDestroy(h->stream);
    delete hh;
    return FDFD_OK;
}

int fdfd_set_grid(fdfd_handle* hh, int Nx, int Ny, double dx, double dy, int npml_x,
                  int npml_y, double L0) {
    Handle* h = H(hh);
    if (!h) return FDFD_ERR_INVALID;
    if (Nx < 1 || Ny < 1 || !(dx > 0) || !(dy > 0) || npml_x < 0 || npml_y < 0 || !(L0 > 0))
        return fail(h, FDFD_ERR_INVALID, "invalid grid parameters");
    if (std::min(Nx, Ny) > MAX_LINE)
        return fail(h, FDFD_ERR_INVALID, "min(Nx, Ny) must be <= 1024");
    (void)hipSetDevice(h->device);
    for (auto& kv : h->factors) free_factor(kv.second);
    h->factors.clear();
    dfree(h->eps);
    h->eps = nullptr;
    h->g = GridDesc{Nx, Ny, dx, dy, npml_x, npml_y, L0};
    h->has_grid = true;
    h->has_eps = false;
    const long N = (long)Nx * Ny;
    HIPCHK(h, hipMalloc(&h->eps, sizeof(double2) * N));
    return FDFD_OK;
}

int fdfd_set_eps_d(fdfd_handle* hh, const void* d_eps) {
    Handle* h = H(hh);
    if (!h || !d_eps) return FDFD_ERR_INVALID;
    if (!h->has_grid) return fail(h, FDFD_ERR_STATE, "fdfd_set_grid must be called first");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    HIPCHK(h, hipMemcpyAsync(h->eps, d_eps, sizeof(double2) * N, hipMemcpyDeviceToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->has_eps = true;
    return FDFD_OK;
}

int fdfd_set_eps(fdfd_handle* hh, const fdfd_complex* eps) {
    Handle* h = H(hh);
    if (!h || !eps) return FDFD_ERR_INVALID;
    if (!h->has_grid) return fail(h, FDFD_ERR_STATE, "fdfd_set_grid must be called first");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    COPY(h, h->eps, eps, sizeof(double2) * N, hipMemcpyHostToDevice);
    h->has_eps = true;
    return FDFD_OK;
}

int fdfd_factor(fdfd_handle* hh, int pol, double omega, int precision, int* factor_id,
                fdfd_stats* stats) {
    Handle* h = H(hh);
    if (!h || !factor_id) return FDFD_ERR_INVALID;
    if (!h->has_grid || !h->has_eps)
        return fail(h, FDFD_ERR_STATE, "grid and eps must be set before fdfd_factor");
    if (pol != FDFD_TM && pol != FDFD_TE) return fail(h, FDFD_ERR_INVALID, "pol must be TM or TE");
    if (precision != FDFD_FP64 && precision != FDFD_FP32_IR)
        return fail(h, FDFD_ERR_INVALID, "precision must be FDFD_FP64 or FDFD_FP32_IR");
    if (!(omega > 0) || !std::isfinite(omega)) return fail(h, FDFD_ERR_INVALID, "omega must be > 0");
    (void)hipSetDevice(h->device);
    Factor* F = new Factor();
    F->pol = pol;
    F->omega = omega;
    F->precision = precision;
    int rc = build_factor(h, F);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->t_factor_ms = F->t_factor_ms;
        stats->pivot_growth = F->pivot_growth;
        stats->singular_line = F->singular_line;
        stats->pivot_fallback = F->pivot_fallback;
    }
    if (rc) {
        free_factor(F);
        return rc;
    }
    F->id = h->next_id++;
    h->factors[F->id] = F;
    *factor_id = F->id;
    return FDFD_OK;
}

int fdfd_release_factor(fdfd_handle* hh, int factor_id) {
    Handle* h = H(hh);
    if (!h) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    h->factors.erase(factor_id);
    free_factor(F);
    return FDFD_OK;
}

static const int MAX_RESTARTS = 40;

int fdfd_solve_d(fdfd_handle* hh, int factor_id, int nrhs, const void* d_src, void* d_out,
                 double tol, fdfd_stats* stats) {
    Handle* h = H(hh);
    if (!h || !d_src || !d_out || nrhs < 1) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    if (!(tol > 0)) return fail(h, FDFD_ERR_INVALID, "tol must be > 0");
    (void)hipSetDevice(h->device);
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
    SolveStats st;
    double worst = 0;
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    HIPCHK(h, hipEventRecord(e0, h->stream));
    int rc = FDFD_OK;
    for (int r = 0; r < nrhs && rc == FDFD_OK; ++r) {
        const double2* src = (const double2*)d_src + (long)r * N;
        double2* out = (double2*)d_out + (long)r * 3 * N;
        // b' = scale * (i omega) * src   (driven.jl:36,46)
        launch_scale_copy(h->stream, h->bvec, src, make_double2(0.0, F->scale * F->omega), N);
        launch_zero(h->stream, h->xvec, N);
        SolveStats s1;
        rc = gmres(h, F, 0, h->bvec, h->xvec, tol, MAX_RESTARTS, s1);
        st.iters += s1.iters;
        st.restarts += s1.restarts;
        worst = std::max(worst, s1.resid);
        if (rc) break;
        // field recovery: TM w = 1/(i omega mu0'), TE w = 1/(i omega eps0')
        double c = (F->pol == FDFD_TM) ? mu0n : eps0n;
        double2 w = make_double2(0.0, -1.0 / (F->omega * c));
        launch_fields(h->stream, out, h->xvec, h->eps, F->sxf, F->sxb, F->syf, F->syb, g, F->pol,
                      w, 0);
    }
    (void)hipEventRecord(e1, h->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iterations = st.iters;
        stats->restarts = st.restarts;
        stats->residual = worst;
        stats->t_factor_ms = F->t_factor_ms;
        stats->t_solve_ms = ms;
        stats->pivot_growth = F->pivot_growth;
        stats->singular_line = F->singular_line;
    }
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return fail(h, FDFD_ERR_HIP, "kernel launch failed");
    return FDFD_OK;
}

int fdfd_solve(fdfd_handle* hh, int factor_id, int nrhs, const fdfd_complex* src,
               fdfd_complex* out, double tol, fdfd_stats* stats) {
    Handle* h = H(hh);
    if (!h || !src || !out || nrhs < 1) return FDFD_ERR_INVALID;
    if (!find_factor(h, factor_id)) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    void *ds = nullptr, *dout = nullptr;
    HIPCHK(h, hipMalloc(&ds, sizeof(double2) * N * nrhs));
    if (hipMalloc(&dout, sizeof(double2) * N * 3 * nrhs) != hipSuccess) {
        dfree(ds);
        return fail(h, FDFD_ERR_OOM, "output allocation failed");
    }
    int rc = FDFD_OK;
    rc = stream_copy(h, ds, src, sizeof(double2) * N * nrhs, hipMemcpyHostToDevice);
    if (!rc) rc = fdfd_solve_d(hh, factor_id, nrhs, ds, dout, tol, stats);
    if (!rc) rc = stream_copy(h, out, dout, sizeof(double2) * N * 3 * nrhs, hipMemcpyDeviceToHost);
    dfree(ds);
    dfree(dout);
    return rc;
}

int fdfd_linsolve(fdfd_handle* hh, int factor_id, int transpose, int nrhs, const fdfd_complex* b,
                  fdfd_complex* x, double tol, fdfd_stats* stats) {
    Handle* h = H(hh);
    if (!h || !b || !x || nrhs < 1) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    if (!(tol > 0)) return fail(h, FDFD_ERR_INVALID, "tol must be > 0");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    SolveStats st;
    double worst = 0;
    for (int r = 0; r < nrhs; ++r) {
        COPY(h, h->tmp, b + (long)r * N, sizeof(double2) * N, hipMemcpyHostToDevice);
        launch_scale_copy(h->stream, h->bvec, h->tmp, make_double2(F->scale, 0), N);
        launch_zero(h->stream, h->xvec, N);
        SolveStats s1;
        int rc = gmres(h, F, transpose ? 1 : 0, h->bvec, h->xvec, tol, MAX_RESTARTS, s1);
        st.iters += s1.iters;
        worst = std::max(worst, s1.resid);
        if (rc) return rc;
        HIPCHK(h, hipMemcpyAsync(x + (long)r * N, h->xvec, sizeof(double2) * N,
                                 hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iterations = st.iters;
        stats->residual = worst;
        stats->t_factor_ms = F->t_factor_ms;
        stats->pivot_growth = F->pivot_growth;
        stats->singular_line = -1;
    }
    return FDFD_OK;
}

int fdfd_apply(fdfd_handle* hh, int factor_id, int transpose, const fdfd_complex* x,
               fdfd_complex* y) {
    Handle* h = H(hh);
    if (!h || !x || !y) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    COPY(h, h->tmp, x, sizeof(double2) * N, hipMemcpyHostToDevice);
    launch_apply(h->stream, h->tmp2, h->tmp, F->coef, h->g.Nx, h->g.Ny, transpose ? 1 : 0,
                 nullptr);
    launch_scale_copy(h->stream, h->tmp2, h->tmp2, make_double2(1.0 / F->scale, 0), N);
    HIPCHK(h, hipMemcpyAsync(y, h->tmp2, sizeof(double2) * N, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FDFD_OK;
}

int fdfd_residual(fdfd_handle* hh, int factor_id, const fdfd_complex* x, const fdfd_complex* b,
                  double* rel) {
    Handle* h = H(hh);
    if (!h || !x || !b || !rel) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    COPY(h, h->tmp, x, sizeof(double2) * N, hipMemcpyHostToDevice);
    COPY(h, h->tmp2, b, sizeof(double2) * N, hipMemcpyHostToDevice);
    launch_scale_copy(h->stream, h->tmp2, h->tmp2, make_double2(F->scale, 0), N);
    double bn = 0, rn = 0;
    int rc = dev_norm(h, h->tmp2, N, &bn);
    if (rc) return rc;
    launch_apply(h->stream, h->bvec, h->tmp, F->coef, h->g.Nx, h->g.Ny, 0, h->tmp2);
    if ((rc = dev_norm(h, h->bvec, N, &rn))) return rc;
    *rel = (bn > 0) ? rn / bn : rn;
    return FDFD_OK;
}

int fdfd_adjoint_grad(fdfd_handle* hh, int factor_id, const fdfd_complex* e_fwd,
                      const fdfd_complex* lam, fdfd_complex weight, double* grad) {
    Handle* h = H(hh);
    if (!h || !e_fwd || !lam || !grad) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    if (F->pol != FDFD_TM)
        return fail(h, FDFD_ERR_INVALID, "adjoint gradient is implemented for TM");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    const double eps0n = EPS0 * h->g.L0;
    COPY(h, h->tmp, e_fwd, sizeof(double2) * N, hipMemcpyHostToDevice);
    COPY(h, h->tmp2, lam, sizeof(double2) * N, hipMemcpyHostToDevice);
    double* dg = (double*)h->bvec;
    const double s = -F->omega * F->omega * eps0n;
    launch_grad(h->stream, dg, h->tmp, h->tmp2, make_double2(s * weight.re, s * weight.im), N);
    HIPCHK(h, hipMemcpyAsync(grad, dg, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FDFD_OK;
}

int fdfd_profile_enable(fdfd_handle* hh, int enable) {
    Handle* h = H(hh);
    if (!h) return FDFD_ERR_INVALID;
    (void)hipSetDevice(h->device);
    prof_resolve(h->prof);
    if (enable) {
        for (int i = 0; i < P_COUNT; ++i) {
            h->prof.ms[i] = 0;
            h->prof.launches[i] = 0;
            h->prof.flops[i] = 0;
            h->prof.bytes[i] = 0;
        }
    }
    h->prof.enabled = enable != 0;
    return FDFD_OK;
}

int fdfd_profile_read(fdfd_handle* hh, int kernel, double* total_ms, long long* launches,
                      double* flops, double* bytes) {
    Handle* h = H(hh);
    if (!h || kernel < 0 || kernel >= P_COUNT) return FDFD_ERR_INVALID;
    (void)hipSetDevice(h->device);
    prof_resolve(h->prof);
    if (total_ms) *total_ms = h->prof.ms[kernel];
    if (launches) *launches = h->prof.launches[kernel];
    if (flops) *flops = h->prof.flops[kernel];
    if (bytes) *bytes = h->prof.bytes[kernel];
    return FDFD_OK;
}

int fdfd_set_tuning(fdfd_handle* hh, int panel_width, int force_full_search) {
    Handle* h = H(hh);
    if (!h) return FDFD_ERR_INVALID;
    if (panel_width < 0 || panel_width > PMAX)
        return fail(h, FDFD_ERR_INVALID, "panel width must be in [0, 32] (0 = default)");
    // force the full-column pivot search on every panel
    h->full_pivot_search = force_full_search != 0;
    h->panel_width = panel_width;
    return FDFD_OK;
}

}  // extern "C"

// exported for eigen.cpp / batch.cpp
namespace fdfd {
int rt_gmres(Handle* h, Factor* F, int trans, const double2* b, double2* x, double tol, int* iters,
             double* resid) {
    SolveStats st;
    int rc = gmres(h, F, trans, b, x, tol, MAX_RESTARTS, st);
    if (iters) *iters += st.iters;
    if (resid) *resid = st.resid;
    return rc;
}
int rt_build_factor(Handle* h, Factor* F) { return build_factor(h, F); }
int rt_build_factors(Handle* h, const std::vector<Factor*>& Fs) { return build_factors(h, Fs); }
void rt_free_factor(Factor* F) { free_factor(F); }
int rt_norm(Handle* h, const double2* x, long N, double* out) { return dev_norm(h, x, N, out); }
int rt_multidot(Handle* h, const double2* V, long N, int k, const double2* w, cd* out) {
    return dev_multidot(h, V, N, k, w, out);
}
int rt_update(Handle* h, double2* w, const double2* V, long N, int k, const cd* c, double sign) {
    return dev_update(h, w, V, N, k, c, sign);
}
}  // namespace fdfd
