// Host runtime and C ABI of libfdfd: handle lifecycle, operator assembly,
// FP64 flexible GMRES refinement, field recovery, adjoint gradient.  The
// block cyclic reduction factorization and its solve are in bcr.cpp,
// shift-invert Arnoldi in eigen.cpp, batched solves in batch.cpp.
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


int stream_copy(Handle* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, kind, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return FDFD_OK;
}


// ---------------------------------------------------------------------------
// profiling: HIP events around launches on the handle's stream
// ---------------------------------------------------------------------------
hipEvent_t ev_get(Prof& p) {
    if (!p.pool.empty()) {
        hipEvent_t e = p.pool.back();
        p.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}
void prof_resolve(Prof& p) {
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

// ---------------------------------------------------------------------------
// work buffers
// ---------------------------------------------------------------------------
static void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int ensure_factor_work(Handle* h, int n, int nch) {
    if (h->work_n >= n && h->work_ch >= nch) return FDFD_OK;
    n = std::max(n, h->work_n);
    nch = std::max(nch, h->work_ch);
    dfree(h->used); dfree(h->piv); dfree(h->Vbuf); dfree(h->LU); dfree(h->Zb);
    dfree(h->win); dfree(h->verr);
    h->used = h->piv = h->win = h->verr = nullptr;
    h->Vbuf = h->LU = h->Zb = nullptr;
    const int NBmax = VMAX;
    HIPCHK(h, hipMalloc(&h->used, sizeof(int) * nch * n));
    HIPCHK(h, hipMalloc(&h->piv, sizeof(int) * nch * NBmax));
    HIPCHK(h, hipMalloc(&h->win, sizeof(int) * nch * WIN));
    HIPCHK(h, hipMalloc(&h->verr, sizeof(int) * MAX_BUILD));  // one flag per factor built together
    HIPCHK(h, hipMalloc(&h->Vbuf, sizeof(double2) * nch * (size_t)n * NBmax));
    HIPCHK(h, hipMalloc(&h->LU, sizeof(double2) * nch * 2 * PMAX * PMAX));
    HIPCHK(h, hipMalloc(&h->Zb, sizeof(double2) * nch * (size_t)NBmax * n));
    h->work_n = n;
    h->work_ch = nch;
    return FDFD_OK;
}

static const int KDIM = GMRES_KDIM;

int ensure_solve_work(Handle* h, long N) {
    if (h->work_N >= N) return FDFD_OK;
    void** bufs[] = {(void**)&h->g_mid, (void**)&h->x_out, (void**)&h->V,
                     (void**)&h->Zk, (void**)&h->tmp, (void**)&h->tmp2, (void**)&h->bvec,
                     (void**)&h->xvec};
    for (auto b : bufs) { dfree(*b); *b = nullptr; }
    HIPCHK(h, hipMalloc(&h->g_mid, sizeof(double2) * N * KMULTI));
    HIPCHK(h, hipMalloc(&h->x_out, sizeof(double2) * N * KMULTI));
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
    }
    h->work_N = N;
    h->kdim = KDIM;
    return FDFD_OK;
}

void* arena_take(Handle* h, size_t bytes) {
    if (h->slab) {  // batch mode: carve from the slab while it lasts
        const size_t off = (h->slab_used + 255) & ~(size_t)255;
        if (off + bytes <= h->slab_size) {
            h->slab_used = off + bytes;
            return h->slab + off;
        }
    }
    int best = -1;
    for (int i = 0; i < (int)h->arenas.size(); ++i)
        if (h->arenas[i].second >= bytes && (best < 0 || h->arenas[i].second < h->arenas[best].second))
            best = i;
    if (best >= 0) {
        void* p = h->arenas[best].first;
        h->arenas.erase(h->arenas.begin() + best);
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {  // retry once without the pooled arenas
        (void)hipGetLastError();
        arena_clear(h);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    return p;
}

void arena_trim(Handle* h) {
    while ((int)h->arenas.size() > h->arena_cap) {  // drop the smallest
        int sm = 0;
        for (int i = 1; i < (int)h->arenas.size(); ++i)
            if (h->arenas[i].second < h->arenas[sm].second) sm = i;
        (void)hipFree(h->arenas[sm].first);
        h->arenas.erase(h->arenas.begin() + sm);
    }
}

size_t arena_pooled_bytes(const Handle* h) {
    size_t t = 0;
    for (const auto& a : h->arenas) t += a.second;
    return t;
}

void arena_give(Handle* h, void* p, size_t bytes) {
    if (!p) return;
    if (h->slab && (char*)p >= h->slab && (char*)p < h->slab + h->slab_size) return;  // slab: reset per group
    h->arenas.push_back({p, bytes});
    arena_trim(h);
}

void arena_clear(Handle* h) {
    for (auto& a : h->arenas) (void)hipFree(a.first);
    h->arenas.clear();
}

void* buf_alloc(Handle* h, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    auto it = h->bpool.find(bytes);
    void* p = nullptr;
    if (it != h->bpool.end()) {
        p = it->second;
        h->bpool.erase(it);
        h->bpool_bytes -= bytes;
    } else if (hipMalloc(&p, bytes) != hipSuccess) {  // retry once without the pools
        (void)hipGetLastError();
        buf_clear(h);
        arena_clear(h);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    h->bpool_live[p] = bytes;
    return p;
}

void buf_free(Handle* h, void* p) {
    if (!p) return;
    auto it = h->bpool_live.find(p);
    if (it == h->bpool_live.end()) {
        // another handle's buffer (a factor a batch's builder handle made)
        (void)hipFree(p);
        return;
    }
    const size_t bytes = it->second;
    h->bpool_live.erase(it);
    if (h->bpool_bytes + bytes > BUF_POOL_MAX) {
        (void)hipFree(p);
        return;
    }
    h->bpool.emplace(bytes, p);
    h->bpool_bytes += bytes;
}

void buf_clear(Handle* h) {
    for (auto& kv : h->bpool) (void)hipFree(kv.second);
    h->bpool.clear();
    h->bpool_bytes = 0;
}

void rt_release_factor_storage(Handle* h, Factor* F) {
    if (!F) return;
    (void)hipStreamSynchronize(h->stream);  // the arena may be reused by the next factor
    for (void* b : {(void*)F->coef, (void*)F->sxf, (void*)F->sxb, (void*)F->syf, (void*)F->syb,
                    (void*)F->prow, (void*)F->iprow, (void*)F->pivstat, F->tabs, (void*)F->triw})
        buf_free(h, b);
    arena_give(h, F->blocks, F->blocks_bytes);
    arena_give(h, F->blocks_b, F->blocks_b_bytes);
    if (F->solve_exec) (void)hipGraphExecDestroy(F->solve_exec);
    F->coef = F->sxf = F->sxb = F->syf = F->syb = nullptr;
    F->blocks = F->blocks_b = nullptr;
    F->prow = F->iprow = nullptr;
    F->pivstat = nullptr;
    F->tabs = nullptr;
    F->mv_tab = nullptr;
    F->dg_tab = nullptr;
    F->trs_tab = nullptr;
    F->triw = nullptr;
    F->solve_exec = nullptr;
}

static void free_factor(Handle* h, Factor* F) {
    if (!F) return;
    rt_release_factor_storage(h, F);
    delete F;
}

// ---------------------------------------------------------------------------
// operator assembly
// ---------------------------------------------------------------------------
int rt_assemble(Handle* h, Factor* F) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
    const double eta0 = std::sqrt(MU0 / EPS0);
    F->coef = (double2*)buf_alloc(h, sizeof(double2) * 5 * N);
    F->sxf = (double2*)buf_alloc(h, sizeof(double2) * g.Nx);
    F->sxb = (double2*)buf_alloc(h, sizeof(double2) * g.Nx);
    F->syf = (double2*)buf_alloc(h, sizeof(double2) * g.Ny);
    F->syb = (double2*)buf_alloc(h, sizeof(double2) * g.Ny);
    if (!F->coef || !F->sxf || !F->sxb || !F->syf || !F->syb)
        return fail(h, FDFD_ERR_OOM, "operator coefficient storage does not fit in device memory");
    if (F->form == 1 && F->pol != FDFD_TM)
        return fail(h, FDFD_ERR_INVALID, "the modulated (sideband) operator is TM only");
    if (F->form == 2 || F->form == 3) {
        // a block of the full-vector mode operator: no stretching (unit
        // factors, so transposed solves and field helpers stay defined)
        const GridDesc g0{g.Nx, g.Ny, g.dx, g.dy, 0, 0, g.L0};
        launch_sfactor(h->stream, F->sxf, F->sxb, F->syf, F->syb, g0, F->omega, eps0n, eta0);
        F->scale = g.dx * g.dx;
        const double k0sq = F->omega * F->omega * mu0n * eps0n;
        ProfScope ps(h, P_ASSEMBLE, 0, 96.0 * N);
        launch_vmode_coeffs(h->stream, F->coef, h->eps, g, k0sq, F->shift, F->scale, F->form - 2);
        HIPCHK(h, hipGetLastError());
        return FDFD_OK;
    }
    launch_sfactor(h->stream, F->sxf, F->sxb, F->syf, F->syb, g,
                   F->omega_pml > 0 ? F->omega_pml : F->omega, eps0n, eta0);
    double kdiag;
    if (F->pol == FDFD_TM) {
        F->scale = mu0n * g.dx * g.dx;
        kdiag = F->omega * F->omega * eps0n * mu0n * g.dx * g.dx;
    } else {
        F->scale = eps0n * g.dx * g.dx;
        kdiag = F->omega * F->omega * mu0n * eps0n * g.dx * g.dx;
    }
    F->kdiag = kdiag;
    {
        // reads eps (16 B/cell), writes the five coefficient planes (80 B/cell)
        ProfScope ps(h, P_ASSEMBLE, 0, 96.0 * N);
        launch_coeffs(h->stream, F->coef, h->eps, F->sxf, F->sxb, F->syf, F->syb, g, F->pol,
                      kdiag, F->form);
    }
    HIPCHK(h, hipGetLastError());
    return FDFD_OK;
}


void rt_apply(Handle* h, const Factor* F, double2* y, const double2* x, int transpose,
              const double2* b) {
    const GridDesc& g = h->g;
    if (F->form == 0 || F->form == 1) {
        const StencilOp op{F->pol, F->form, g.Nx, g.Ny, h->eps, F->sxf, F->sxb, F->syf, F->syb,
                           (g.dx * g.dx) / (g.dy * g.dy), F->kdiag};
        launch_stencil(h->stream, y, x, op, transpose, b);
    } else {
        launch_apply(h->stream, y, x, F->coef, g.Nx, g.Ny, transpose, b);
    }
}

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
int rt_gmres_op(Handle* h, long N,
                const std::function<int(const double2*, double2*, const double2*)>& apply,
                const std::function<int(const double2*, double2*)>& prec, double2* V, double2* Zk,
                const double2* b, double2* x, double tol, int max_restarts, int* iters,
                int* restarts, double* resid_out) {
    int K = GMRES_KDIM;
    double bnorm;
    int rc;
    if (h->inject == FDFD_INJECT_NOCONV) {  // one-shot: at most inject_arg iterations, no restart
        h->inject = 0;
        K = std::min(std::max(h->inject_arg, 0), GMRES_KDIM);
        max_restarts = K > 0 ? 1 : 0;
    }
    if ((rc = dev_norm(h, b, N, &bnorm))) return rc;
    if (bnorm == 0) {
        launch_zero(h->stream, x, N);
        *resid_out = 0;
        return FDFD_OK;
    }
    std::vector<cd> H((size_t)(K + 1) * K), gv(K + 1), sn(K), hc(K + 1), hc2(K + 1), y(K);
    std::vector<double> cs(K);
    double resid = 1, prev = 2;
    for (int rs = 0; rs <= max_restarts; ++rs) {
        if ((rc = apply(x, V, b))) return rc;
        double beta;
        if ((rc = dev_norm(h, V, N, &beta))) return rc;
        resid = beta / bnorm;
        *resid_out = resid;
        if (resid <= tol) return FDFD_OK;
        if (rs == max_restarts) break;
        // a whole restart cycle that gains less than 10% has reached the
        // rounding floor of the operator: stop instead of cycling
        if (rs > 0 && resid > 0.9 * prev) break;
        prev = resid;
        if (rs > 0) ++*restarts;
        launch_scale_copy(h->stream, V, V, make_double2(1.0 / beta, 0), N);
        std::fill(gv.begin(), gv.end(), cd(0));
        gv[0] = beta;
        int jdone = 0;
        for (int j = 0; j < K; ++j) {
            double2* vj = V + (long)j * N;
            double2* zj = Zk + (long)j * N;
            double2* w = V + (long)(j + 1) * N;
            if ((rc = prec(vj, zj))) return rc;
            if ((rc = apply(zj, w, nullptr))) return rc;
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
            ++*iters;
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
    *resid_out = resid;
    char buf[160];
    snprintf(buf, sizeof(buf),
             "refinement did not reach the tolerance %.1e (relative residual %.3e)", tol, resid);
    return fail(h, FDFD_ERR_NOCONV, buf);
}

// FGMRES on one factor's operator A' (or A'^T), preconditioned by the factor
static int gmres(Handle* h, Factor* F, int trans, const double2* b, double2* x, double tol,
                 int max_restarts, SolveStats& st) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    auto apply = [&](const double2* xx, double2* yy, const double2* bb) -> int {
        // x, eps, y (and b) at 16 B per cell
        ProfScope ps(h, P_STENCIL, 0, (bb ? 64.0 : 48.0) * N);
        rt_apply(h, F, yy, xx, trans, bb);
        return FDFD_OK;
    };
    auto prec = [&](const double2* v, double2* z) -> int { return rt_prec_apply(h, F, trans, v, z); };
    return rt_gmres_op(h, N, apply, prec, h->V, h->Zk, b, x, tol, max_restarts, &st.iters,
                       &st.restarts, &st.resid);
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
    // Reduced-solve graphs are replayed unless FDFD_GRAPHS=0, or unless the
    // rocprofv3 tool library is loaded into the process (it exports
    // ROCPROFILER_LIBRARY_CTOR to the profiled program): a modulated solve
    // replaying graphs under rocprofv3's kernel trace ended in a host
    // segmentation fault inside the graph calls, with no host access to
    // device memory in this library's code, so profiled runs use direct
    // launches (the same kernels, in the same order).
    const char* eg = std::getenv("FDFD_GRAPHS");
    const char* prof_tool = std::getenv("ROCPROFILER_LIBRARY_CTOR");
    hh->impl.use_graphs = !(eg && eg[0] == '0') && !(prof_tool && prof_tool[0]);
    const char* e3 = std::getenv("FDFD_GEMM3M");
    hh->impl.gemm3m = (e3 && e3[0] == '0') ? 0 : 1;
    const char* ew = std::getenv("FDFD_GEMM_WAVES");
    hh->impl.gemm_waves = (ew && ew[0] == '4') ? 4 : 8;
    const char* eb = std::getenv("FDFD_MIXED_GEMM");
    hh->impl.gemm_b3 = (eb && std::strcmp(eb, "fp32") == 0) ? 0 : 1;
    const char* et = std::getenv("FDFD_MIXED_TRAIL");
    hh->impl.trail_h3 = (et && std::strcmp(et, "fp32") == 0) ? 0 : 1;
    const char* ef = std::getenv("FDFD_FUSE_PANELS");
    hh->impl.fuse_panels = (ef && ef[0] == '0') ? 0 : 1;
    hh->impl.env_knobs = e3 || ew || eb || et || ef;
    const char* em = std::getenv("FDFD_MV_MFMA");
    hh->impl.mv_mfma = (em && em[0] == '0') ? 0 : 1;
    if (hipStreamCreateWithFlags(&hh->impl.stream, hipStreamNonBlocking) != hipSuccess) {
        delete hh;
        return FDFD_ERR_HIP;
    }
    // per-device static scratch of the no-allocation entry points
    if (rt_flux_scratch_setup(&hh->impl) != FDFD_OK) {
        (void)hipStreamDestroy(hh->impl.stream);
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
    for (auto& kv : h->factors) free_factor(h, kv.second);
    arena_clear(h);
    buf_clear(h);
    free_multi_work(h);
    void* bufs[] = {h->eps, h->used, h->piv, h->win, h->verr, h->Vbuf, h->LU, h->Zb,
                    h->g_mid, h->x_out, h->V, h->Zk, h->tmp, h->tmp2, h->bvec, h->xvec,
                    h->dpart, h->rpart, h->hdev, h->bslab, h->io};
    for (void* b : bufs) dfree(b);
    if (h->tstage) (void)hipHostFree(h->tstage);
    if (h->tstage_ev) (void)hipEventDestroy(h->tstage_ev);
    prof_resolve(h->prof);
    for (auto e : h->prof.pool) (void)hipEventDestroy(e);
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
    for (auto& kv : h->factors) free_factor(h, kv.second);
    h->factors.clear();
    // The pooled factor arenas and the eps buffer are kept when the grid keeps
    // its shape (a new design or frequency sweep on the same grid): freeing
    // and reallocating gigabytes makes the driver clear the memory again,
    // which slows the GPU work that follows.
    const bool same_shape = h->has_grid && h->g.Nx == Nx && h->g.Ny == Ny;
    if (!same_shape) {
        arena_clear(h);
        buf_clear(h);
    }
    h->g = GridDesc{Nx, Ny, dx, dy, npml_x, npml_y, L0};
    h->has_grid = true;
    h->has_eps = false;
    const long N = (long)Nx * Ny;
    if (h->eps_n < N) {
        dfree(h->eps);
        h->eps = nullptr;
        h->eps_n = 0;
        HIPCHK(h, hipMalloc(&h->eps, sizeof(double2) * N));
        h->eps_n = N;
    }
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
    int rc = rt_build_factor(h, F);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->t_factor_ms = F->t_factor_ms;
        stats->pivot_growth = F->pivot_growth;
        stats->singular_line = F->singular_line;
        stats->pivot_fallback = F->pivot_fallback;
    }
    if (rc) {
        free_factor(h, F);
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
    free_factor(h, F);
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
    // field recovery: TM w = 1/(i omega mu0'), TE w = 1/(i omega eps0')
    const double2 wf = make_double2(0.0, -1.0 / (F->omega * ((F->pol == FDFD_TM) ? mu0n : eps0n)));
    if (nrhs > 1) {
        // several sources: lockstep FGMRES, KMULTI at a time, one pass over
        // the factor per iteration for all of them
        for (int s0 = 0; s0 < nrhs && rc == FDFD_OK; s0 += KMULTI) {
            const int ns = std::min(KMULTI, nrhs - s0);
            if ((rc = ensure_multi_work(h, N, ns))) break;
            for (int k = 0; k < ns; ++k) {
                launch_scale_copy(h->stream, h->mB + (long)k * N, (const double2*)d_src + (long)(s0 + k) * N,
                                  make_double2(0.0, F->scale * F->omega), N);
                launch_zero(h->stream, h->mX + (long)k * N, N);
            }
            rc = rt_gmres_multi(h, F, ns, h->mB, h->mX, tol, &st.iters, &worst);
            if (rc) break;
            for (int k = 0; k < ns; ++k)
                launch_fields(h->stream, (double2*)d_out + (long)(s0 + k) * 3 * N, h->mX + (long)k * N,
                              h->eps, F->sxf, F->sxb, F->syf, F->syb, g, F->pol, wf, 0);
        }
    }
    for (int r = 0; r < nrhs && nrhs == 1 && rc == FDFD_OK; ++r) {
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
    // device staging: sources then fields, in the handle's buffer (grown, not
    // reallocated per call)
    const size_t need = sizeof(double2) * N * 4 * (size_t)nrhs;
    if (h->io_bytes < need) {
        dfree(h->io);
        h->io = nullptr;
        h->io_bytes = 0;
        HIPCHK(h, hipMalloc(&h->io, need));
        h->io_bytes = need;
    }
    double2* ds = (double2*)h->io;
    double2* dout = ds + N * nrhs;
    int rc = stream_copy(h, ds, src, sizeof(double2) * N * nrhs, hipMemcpyHostToDevice);
    if (!rc) rc = fdfd_solve_d(hh, factor_id, nrhs, ds, dout, tol, stats);
    if (!rc) rc = stream_copy(h, out, dout, sizeof(double2) * N * 3 * nrhs, hipMemcpyDeviceToHost);
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
    rt_apply(h, F, h->tmp2, h->tmp, transpose ? 1 : 0, nullptr);
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
    rt_apply(h, F, h->bvec, h->tmp, 0, h->tmp2);
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

int fdfd_solve_pullback_d(fdfd_handle* hh, int factor_id, const void* d_ez, const void* d_dfields,
                          double tol, void* d_grad, fdfd_stats* stats) {
    Handle* h = H(hh);
    if (!h || !d_ez || !d_dfields || !d_grad) return FDFD_ERR_INVALID;
    Factor* F = find_factor(h, factor_id);
    if (!F) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    if (F->pol != FDFD_TM || F->form != 0)
        return fail(h, FDFD_ERR_INVALID, "the solve pullback is implemented for TM driven factors");
    if (!(tol > 0)) return fail(h, FDFD_ERR_INVALID, "tol must be > 0");
    (void)hipSetDevice(h->device);
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const double eps0n = EPS0 * g.L0, mu0n = MU0 * g.L0;
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    HIPCHK(h, hipEventRecord(e0, h->stream));
    // scale * conj(bar Ez + Mx^H bar Hx + My^H bar Hy), c = 1/(i w mu0')
    launch_field_adjoint(h->stream, h->bvec, (const double2*)d_dfields, F->sxb, F->syb, g,
                         make_double2(0.0, -1.0 / (F->omega * mu0n)), F->scale);
    launch_zero(h->stream, h->xvec, N);
    SolveStats st;
    int rc = gmres(h, F, 1, h->bvec, h->xvec, tol, MAX_RESTARTS, st);
    if (rc == FDFD_OK) {
        // dL/deps_k = -Re(lam_k w^2 eps0' e_k)   (dA/deps_k = w^2 eps0' e_k e_k^T, driven.jl:21,35)
        launch_grad(h->stream, (double*)d_grad, (const double2*)d_ez, h->xvec,
                    make_double2(-F->omega * F->omega * eps0n, 0.0), N);
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
        stats->residual = st.resid;
        stats->t_factor_ms = F->t_factor_ms;
        stats->t_solve_ms = ms;
        stats->pivot_growth = F->pivot_growth;
        stats->singular_line = -1;
    }
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return fail(h, FDFD_ERR_HIP, "kernel launch failed");
    return FDFD_OK;
}

int fdfd_solve_pullback(fdfd_handle* hh, int factor_id, const fdfd_complex* ez,
                        const fdfd_complex* dfields, double tol, double* grad, fdfd_stats* stats) {
    Handle* h = H(hh);
    if (!h || !ez || !dfields || !grad) return FDFD_ERR_INVALID;
    if (!find_factor(h, factor_id)) return fail(h, FDFD_ERR_NOTFOUND, "unknown factor id");
    (void)hipSetDevice(h->device);
    const long N = (long)h->g.Nx * h->g.Ny;
    void* buf = nullptr;  // ez | dfields (3N) | grad
    HIPCHK(h, hipMalloc(&buf, sizeof(double2) * 4 * N + sizeof(double) * N));
    double2* dez = (double2*)buf;
    double2* ddf = dez + N;
    double* dg = (double*)(ddf + 3 * N);
    int rc = stream_copy(h, dez, ez, sizeof(double2) * N, hipMemcpyHostToDevice);
    if (!rc) rc = stream_copy(h, ddf, dfields, sizeof(double2) * 3 * N, hipMemcpyHostToDevice);
    if (!rc) rc = fdfd_solve_pullback_d(hh, factor_id, dez, ddf, tol, dg, stats);
    if (!rc) rc = stream_copy(h, grad, dg, sizeof(double) * N, hipMemcpyDeviceToHost);
    dfree(buf);
    return rc;
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

int fdfd_debug_inject(fdfd_handle* hh, int what, int arg) {
    Handle* h = H(hh);
    if (!h) return FDFD_ERR_INVALID;
    if (what < FDFD_INJECT_NONE || what > FDFD_INJECT_OOM)
        return fail(h, FDFD_ERR_INVALID, "unknown injection");
    h->inject = what;
    h->inject_arg = arg;
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
void rt_free_factor(Handle* h, Factor* F) { free_factor(h, F); }
int rt_norm(Handle* h, const double2* x, long N, double* out) { return dev_norm(h, x, N, out); }
int rt_multidot(Handle* h, const double2* V, long N, int k, const double2* w, cd* out) {
    return dev_multidot(h, V, N, k, w, out);
}
int rt_update(Handle* h, double2* w, const double2* V, long N, int k, const cd* c, double sign) {
    return dev_update(h, w, V, N, k, c, sign);
}
}  // namespace fdfd
