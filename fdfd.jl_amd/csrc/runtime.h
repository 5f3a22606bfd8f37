// Host runtime structures of libfdfd (C++; no types from here cross the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/fdfd.h"
#include "fdfd_common.h"

namespace fdfd {

using cd = std::complex<double>;

// per-kernel profile slots: Gauss-Jordan trailing update, panel, block
// formation, reduced-solve steps, stencil/residual, Z build, complex GEMM,
// operator assembly
enum ProfId { P_TRAIL = 0, P_PANEL = 1, P_SCHUR = 2, P_SWEEP = 3, P_STENCIL = 4, P_ZBUILD = 5,
              P_GEMM = 6, P_ASSEMBLE = 7, P_COUNT = 8 };

struct Prof {
    bool enabled = false;
    struct Pending { hipEvent_t a, b; int id; double flops, bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double ms[P_COUNT] = {0};
    long long launches[P_COUNT] = {0};
    double flops[P_COUNT] = {0};
    double bytes[P_COUNT] = {0};
};

// Factors of one operator (pol, omega) on the handle's grid: the blocks of
// the block cyclic reduction and its reduced-solve program.
struct Factor {
    int id = 0, pol = 1, precision = 0;
    double omega = 0;
    // operator form: 0 = driven (Sxf dxf Sxb dxb + ..., driven.jl:35),
    // 1 = sideband operator of the modulated solve (Sxb dxb Sxf dxf + ...,
    // modulation.jl:83,88; TM only)
    // 2, 3 = the Hx-Hx / Hy-Hy blocks of the full-vector mode operator minus
    // `shift` (eigen.jl:45; no PML, scale dx^2)
    int form = 0;
    double omega_pml = 0;          // frequency of the stretch factors (0: omega)
    double2 shift = {0, 0};        // sigma of forms 2, 3 (unscaled)
    int n = 0, m = 0, orient = 0;  // line length, number of lines, 0: lines along x
    long si = 1, sj = 1;           // stride of a line's cells / of consecutive lines
    double scale = 1;              // A' = scale * A
    double kdiag = 0;              // diagonal term of A' (times eps for TM), see k_coeffs
    double2* coef = nullptr;       // 5 planes of N
    double2 *sxf = nullptr, *sxb = nullptr, *syf = nullptr, *syb = nullptr;
    size_t elem = 16;              // bytes per complex factor entry
    void* blocks = nullptr;        // nblocks * n * n complex (double2 or float2)
    int nblocks = 0;
    void* blocks_b = nullptr;      // split GEMM operands (fp16 planes + row exponents), ~8 B per entry
    size_t blocks_bytes = 0, blocks_b_bytes = 0;
    int *prow = nullptr, *iprow = nullptr;  // m * n Gauss-Jordan pivot maps per line
    double* pivstat = nullptr;     // 3 per line
    // kind 0: MvJob range, 1: DiagJob range, 2: TrsJob range; mats: n x n
    // blocks the step streams
    struct SolveStep { int kind, first, count, mats; };
    std::vector<SolveStep> ssteps;
    void* tabs = nullptr;  // one buffer holding the three tables below
    MvJob* mv_tab = nullptr;
    DiagJob* dg_tab = nullptr;
    TrsJob* trs_tab = nullptr;
    double* triw = nullptr;        // Givens QR work of the first-level lines (tris order)
    hipGraphExec_t solve_exec = nullptr;
    double t_factor_ms = 0;
    double pivot_growth = 0;
    int singular_line = -1;
    int pivot_fallback = 0;   // 1 if the full-column pivot search was used
};

struct Handle {
    int device = 0;
    hipStream_t stream = nullptr;
    GridDesc g{};
    bool has_grid = false, has_eps = false;
    double2* eps = nullptr;
    std::map<int, Factor*> factors;
    int next_id = 1;
    std::string err;
    int panel_width = 0;  // 0 = default per precision
    // Gauss-Jordan work, one channel per line inverted concurrently
    int work_n = 0, work_ch = 0;
    int *used = nullptr, *piv = nullptr, *win = nullptr, *verr = nullptr;
    void *Vbuf = nullptr, *LU = nullptr, *Zb = nullptr;
    bool full_pivot_search = false;
    // one-shot failure injection (fdfd_debug_inject): FDFD_INJECT_* and its argument
    int inject = 0, inject_arg = 0;
    bool use_graphs = true;            // FDFD_GRAPHS=0 disables graph replay
    int gemm3m = 1;                    // three-multiplication complex GEMM (FDFD_GEMM3M=0: four)
    int mv_mfma = 1;                   // multi-source reduced-solve products on MFMA (FDFD_MV_MFMA=0: VALU)
    int gemm_waves = 8;                // waves per GEMM workgroup (FDFD_GEMM_WAVES=4: four)
    int gemm_b3 = 1;                   // mixed path: three-product fp16 GEMMs (FDFD_MIXED_GEMM=fp32: FP32)
    int trail_h3 = 1;                  // mixed path: three-product fp16 trailing updates (FDFD_MIXED_TRAIL=fp32: FP32)
    int fuse_panels = 1;               // two Gauss-Jordan panels per trailing pass (FDFD_FUSE_PANELS=0)
    bool env_knobs = false;            // a variant switch above came from the environment
    bool tuned_explicitly = false;     // fdfd_autotune chose the variants (no cache lookup)
    // sweep / Krylov work
    long work_N = 0;
    double2 *g_mid = nullptr, *x_out = nullptr;  // reduced-solve vectors (KMULTI sources each)
    // multi-source FGMRES work (multi.cpp): bases for up to multi_k sources of
    // multi_N cells, small per-source device arrays and pinned staging
    long multi_N = 0;
    int multi_k = 0;
    double2 *mV = nullptr, *mZ = nullptr, *mB = nullptr, *mX = nullptr;
    void* msmall = nullptr;
    void* mpin = nullptr;
    void* mpack = nullptr;             // packed inputs of the MFMA reduced-solve products
    size_t mpack_bytes = 0;
    double2 *V = nullptr, *Zk = nullptr;  // Krylov bases
    int kdim = 0;
    double2 *tmp = nullptr, *tmp2 = nullptr, *bvec = nullptr, *xvec = nullptr;
    double2* dpart = nullptr;   // multidot partials
    double* rpart = nullptr;    // norm partials
    double2* hdev = nullptr;    // small device vector of coefficients
    int nblk = 0;
    Prof prof;
    // released factor arenas kept for reuse (at most ARENA_POOL), freed by
    // fdfd_set_grid and fdfd_destroy
    std::vector<std::pair<void*, size_t>> arenas;
    int arena_cap = 2;   // pool capacity (ARENA_POOL; raised while a batch runs)
    // while a batch runs, factor arenas are carved from one slab allocated for
    // the batch's largest group (bump pointer, reset per group)
    char* slab = nullptr;
    size_t slab_size = 0, slab_used = 0;
    // the batch slab stays allocated between batches (freeing and
    // reallocating tens of GB makes the driver clear the memory again, which
    // slows the GPU work that follows for seconds); freed by fdfd_destroy
    char* bslab = nullptr;
    size_t bslab_size = 0;
    // per-factor buffers (coefficient planes, stretch vectors, pivot maps,
    // line-solve work, job tables) kept for reuse by exact size: every
    // factorization of a grid allocates the same set, and hipMalloc/hipFree
    // pairs cost about a millisecond per factorization
    std::multimap<size_t, void*> bpool;      // free buffers by size
    // pinned staging of job-table uploads and the event of the last upload
    void* tstage = nullptr;
    size_t tstage_bytes = 0;
    hipEvent_t tstage_ev = nullptr;
    std::map<void*, size_t> bpool_live;      // handed-out buffers and their sizes
    size_t bpool_bytes = 0;                  // bytes held free in bpool
    // device staging of the host-pointer solve (sources, then fields), kept
    // between calls and grown as needed
    void* io = nullptr;
    size_t io_bytes = 0;
    long eps_n = 0;      // cells the eps buffer holds
};

// error helper: record message, return code
int fail(Handle* h, int code, const std::string& msg);

#define HIPCHK(h, call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail((h), e_ == hipErrorOutOfMemory ? FDFD_ERR_OOM : FDFD_ERR_HIP, \
                        std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define COPY(h, dst, src, bytes, kind)                             \
    do {                                                           \
        int rc_ = stream_copy((h), (dst), (src), (bytes), (kind)); \
        if (rc_) return rc_;                                       \
    } while (0)

// Copy ordered on the handle's stream, complete on return.  The handle's
// streams are non-blocking, so a plain hipMemcpy (null stream) would not wait
// for kernels still queued on them; every host<->device transfer of the
// library goes through here instead.
int stream_copy(Handle* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);


// HIP events around launches on a stream while profiling is enabled
hipEvent_t ev_get(Prof& p);
void prof_resolve(Prof& p);
struct ProfScope {
    Handle* h;
    int id;
    double flops, bytes;
    hipStream_t st;
    hipEvent_t a{}, b{};
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

// factor arenas: take one of at least `bytes` from the pool or allocate;
// give one back (the pool keeps the largest ARENA_POOL); free them all
constexpr int ARENA_POOL = 2;
constexpr int MAX_BUILD = 128;   // factors one rt_build_factors call builds together
constexpr int KMULTI = 16;       // sources one multi-source FGMRES solves together
void arena_trim(Handle* h);      // drop the smallest arenas beyond arena_cap
size_t arena_pooled_bytes(const Handle* h);
void* arena_take(Handle* h, size_t bytes);
void arena_give(Handle* h, void* p, size_t bytes);
void arena_clear(Handle* h);
// pooled per-factor buffers: buf_alloc returns nullptr when device memory is
// short; buf_free takes nullptr and pointers from buf_alloc (of this handle,
// or of another, whose buffer is then freed); buf_clear frees the free list
// (BUF_POOL_MAX bytes at most are kept)
constexpr size_t BUF_POOL_MAX = (size_t)1 << 30;
void* buf_alloc(Handle* h, size_t bytes);
void buf_free(Handle* h, void* p);
void buf_clear(Handle* h);
void rt_free_factor(Handle* h, Factor* F);
void rt_release_factor_storage(Handle* h, Factor* F);  // device storage only, F stays
// FGMRES for nr <= KMULTI sources on one factor in lockstep (multi.cpp):
// X_r solves A' X_r = B_r (vectors at r * N), each to ||B_r - A' X_r|| <= tol ||B_r||
int rt_gmres_multi(Handle* h, Factor* F, int nr, const double2* B, double2* X, double tol,
                   int* iters, double* worst);
int ensure_multi_work(Handle* h, long N, int k);
// variants from the FDFD_TUNE_CACHE file for this shape (tune.cpp)
void rt_tune_lookup(Handle* h, int n, int m, int precision, int pol);
void free_multi_work(Handle* h);

// work buffers of the handle (grown on demand)
int ensure_factor_work(Handle* h, int n, int nch);
int ensure_solve_work(Handle* h, long N);
int rt_flux_scratch_setup(Handle* h);
// y = A' x, A'^T x (transpose) or b - A' x (b != nullptr) for factor F's
// operator: the matrix-free stencil for driven and sideband operators, the
// coefficient planes for the full-vector mode blocks
void rt_apply(Handle* h, const Factor* F, double2* y, const double2* x, int transpose,
              const double2* b);

// Right-preconditioned FGMRES(30) in FP64 on an operator given by callbacks:
//   apply(x, y, b): y = b - A x if b != nullptr, else y = A x
//   prec(v, z):     z = M^-1 v
// V ((30 + 1) * N) and Zk (30 * N) are the caller's device workspace.  Stops
// at ||b - A x|| <= tol ||b|| (FDFD_ERR_NOCONV after max_restarts restarts).
constexpr int GMRES_KDIM = 30;
int rt_gmres_op(Handle* h, long N, const std::function<int(const double2*, double2*, const double2*)>& apply,
                const std::function<int(const double2*, double2*)>& prec, double2* V, double2* Zk,
                const double2* b, double2* x, double tol, int max_restarts, int* iters,
                int* restarts, double* resid);

// block cyclic reduction (bcr.cpp)
int rt_build_factors(Handle* h, const std::vector<Factor*>& Fs);
int rt_build_factor(Handle* h, Factor* F);
int rt_prec_apply(Handle* h, Factor* F, int trans, const double2* in, double2* out);
size_t rt_factor_bytes(int n, int m, int precision, bool b3);
void rt_factor_arena_sizes(int n, int m, int precision, bool b3, size_t* blocks, size_t* split);

}  // namespace fdfd

struct fdfd_handle { fdfd::Handle impl; };
