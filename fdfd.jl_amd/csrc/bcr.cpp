// Block cyclic reduction of the line-block operator (host side).
//
// Lines j = 0..m-1 (n cells each) form a ring: line j couples to j-1 and j+1
// through diagonal blocks, and the y-periodic corner blocks close the ring
// (src/grid.jl:138-150 wraps the differences).  A reduction level removes
// every other line of the current ring (all of them at once):
//   X_j = D_j^-1                                   (Gauss-Jordan, batched)
//   D_s -= C_sj X_j C_js          for both surviving neighbours s of j
//   C'(s, u) = [C(s, u)] - C_sj X_j C_ju   (new coupling s -> u across j)
// until one line is left.  The critical path is log2(m) batched inversions
// instead of the m sequential ones of a block-Thomas chain, and the corner
// blocks are kept exactly, so the factorization is a direct solver of A'.
//
// The plan (which blocks, which jobs, in which order) depends on m only; it
// is built symbolically and then instantiated for each factor as device job
// tables.  Block ids 0..m-1 are the lines' own blocks (the inverse of line j
// stays in block j); coupling and product blocks are drawn from a pool.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/fdfd.h"
#include "bcr_plan.h"
#include "runtime.h"
#include "tri_qr.h"

namespace fdfd {

using namespace bcr;

int rt_assemble(Handle* h, Factor* F);

// The reduction plan depends on m only and never changes: built once per m
// and shared (a batch's pipeline builds factors from two threads).
static const Plan& plan_for(int m) {
    static std::mutex mu;
    static std::map<int, std::unique_ptr<const Plan>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto& p = cache[m];
    if (!p) p.reset(new Plan(make_plan(m)));
    return *p;
}

size_t rt_factor_bytes(int n, int m, int precision, bool b3) {
    const size_t elem = (precision == FDFD_FP64) ? sizeof(double2) : sizeof(float2);
    const Plan& P = plan_for(m);
    size_t bytes = (size_t)P.nblocks * n * n * elem + sizeof(int) * 2 * (size_t)m * n;
    // Givens QR work of the first-level lines, kept for the reduced solve
    bytes += P.tris.size() * (size_t)tri::work_doubles(n) * sizeof(double);
    // three-product GEMM operands: four tiled fp16 planes and exponents per block
    if (b3 && precision != FDFD_FP64) bytes += (size_t)P.nblocks * b3_block_bytes(n);
    return bytes;
}

// the two arenas a factor takes: its blocks and (fp16x3 GEMMs) the split operands
void rt_factor_arena_sizes(int n, int m, int precision, bool b3, size_t* blocks, size_t* split) {
    const size_t elem = (precision == FDFD_FP64) ? sizeof(double2) : sizeof(float2);
    const Plan& P = plan_for(m);
    *blocks = (size_t)P.nblocks * n * n * elem;
    *split = (b3 && precision != FDFD_FP64) ? (size_t)P.nblocks * b3_block_bytes(n) : 0;
}

// ---------------------------------------------------------------------------
// device tables for a set of factors built together (same grid and shape)
// ---------------------------------------------------------------------------
struct Tables {
    Handle* h;
    void* base = nullptr;  // one buffer holds every table
    SplitJob* split = nullptr;
    TriJob* tri = nullptr;
    FormJob* form = nullptr;
    LineJob* lj = nullptr;
    GemmJob* gemm = nullptr;
    explicit Tables(Handle* hh) : h(hh) {}
    ~Tables() { buf_free(h, base); }
};

// Several host job tables packed into one device buffer (256-byte aligned
// parts) and uploaded with one copy: add() the tables, then commit().
struct TablePack {
    std::vector<char> bytes;
    std::vector<std::pair<void**, size_t>> parts;  // destination pointer, offset
    template <typename T> void add(T** dst, const std::vector<T>& v) {
        *dst = nullptr;
        if (v.empty()) return;
        const size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off + sizeof(T) * v.size());
        std::memcpy(bytes.data() + off, v.data(), sizeof(T) * v.size());
        parts.push_back({(void**)dst, off});
    }
    // *base receives the buffer (nullptr when every table is empty).  The
    // copy goes through the handle's pinned staging buffer and is ordered on
    // the stream before the kernels that read the tables, so the host does
    // not wait for it (only the next upload waits, for the staging buffer).
    int commit(Handle* h, void** base) {
        *base = nullptr;
        if (bytes.empty()) return FDFD_OK;
        void* p = buf_alloc(h, bytes.size());
        if (!p) return fail(h, FDFD_ERR_OOM, "job table storage does not fit in device memory");
        *base = p;
        for (auto& q : parts) *q.first = (char*)p + q.second;
        if (h->tstage_ev) HIPCHK(h, hipEventSynchronize(h->tstage_ev));  // the previous upload is done
        if (h->tstage_bytes < bytes.size()) {
            if (h->tstage) (void)hipHostFree(h->tstage);
            h->tstage = nullptr;
            h->tstage_bytes = 0;
            HIPCHK(h, hipHostMalloc(&h->tstage, bytes.size()));
            h->tstage_bytes = bytes.size();
        }
        if (!h->tstage_ev) HIPCHK(h, hipEventCreateWithFlags(&h->tstage_ev, hipEventDisableTiming));
        std::memcpy(h->tstage, bytes.data(), bytes.size());
        HIPCHK(h, hipMemcpyAsync(p, h->tstage, bytes.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipEventRecord(h->tstage_ev, h->stream));
        return FDFD_OK;
    }
};

static const double2* plane_ptr(const Factor* F, long N, int line, int pl) {
    // plane 0: coupling to the previous line, plane 1: to the next line
    const double2* c = F->coef;
    int k;
    if (F->orient == 0) k = pl ? CN : CS;
    else k = pl ? CE : CW;
    return c + (long)k * N + (long)line * F->sj;
}

static void own_planes(const Factor* F, long N, int line, const double2** sub,
                       const double2** diag, const double2** sup) {
    const double2* c = F->coef;
    const long off = (long)line * F->sj;
    if (F->orient == 0) {
        *sub = c + (long)CW * N + off;
        *sup = c + (long)CE * N + off;
    } else {
        *sub = c + (long)CS * N + off;
        *sup = c + (long)CN * N + off;
    }
    *diag = c + (long)CC * N + off;
}

static void* blk(const Factor* F, int id) {
    return (char*)F->blocks + (size_t)id * F->n * F->n * F->elem;
}
// split operand of block id (three-product GEMMs): 4 tiled fp16 planes, then
// the row exponents
static void* blk_b(const Factor* F, int id) {
    return (char*)F->blocks_b + (size_t)id * b3_block_bytes(F->n);
}
static int* blk_e(const Factor* F, int id) {
    return (int*)((char*)blk_b(F, id) + 4 * (size_t)b3_plane(F->n) * 2);
}

int rt_build_factors(Handle* h, const std::vector<Factor*>& Fs) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const int B = (int)Fs.size();
    if (B < 1) return fail(h, FDFD_ERR_INVALID, "no factor to build");
    if (B > MAX_BUILD)
        return fail(h, FDFD_ERR_INVALID, "at most " + std::to_string(MAX_BUILD) + " factors per build");
    const int prec = Fs[0]->precision;
    for (Factor* F : Fs) {
        if (F->precision != prec)
            return fail(h, FDFD_ERR_INVALID, "factors built together share one precision");
        if (g.Nx <= g.Ny) { F->orient = 0; F->n = g.Nx; F->m = g.Ny; F->si = 1; F->sj = g.Nx; }
        else { F->orient = 1; F->n = g.Ny; F->m = g.Nx; F->si = g.Nx; F->sj = 1; }
        F->elem = (prec == FDFD_FP64) ? sizeof(double2) : sizeof(float2);
    }
    const int n = Fs[0]->n, m = Fs[0]->m;
    const size_t elem = Fs[0]->elem;
    rt_tune_lookup(h, n, m, prec, Fs[0]->pol);
    // mixed path with three-product GEMMs: operands are split once per step
    const bool b3 = (prec != FDFD_FP64) && h->gemm_b3;
    if (n > MAX_LINE)
        return fail(h, FDFD_ERR_INVALID, "line block larger than 1024 cells (min(Nx,Ny) > 1024)");
    const Plan& P = plan_for(m);
    int rc;
    const int inject = h->inject;  // one-shot: consumed by this factorization
    if (inject == FDFD_INJECT_PIVOT_FALLBACK || inject == FDFD_INJECT_SINGULAR ||
        inject == FDFD_INJECT_OOM)
        h->inject = 0;
    if (inject == FDFD_INJECT_OOM)
        return fail(h, FDFD_ERR_OOM, "factor storage allocation failed (injected)");
    if ((rc = ensure_factor_work(h, n, std::max(1, P.max_gj * B)))) return rc;
    if ((rc = ensure_solve_work(h, N))) return rc;
    if (trisolve_prepare(n) != (int)hipSuccess)
        return fail(h, FDFD_ERR_HIP, "line-solve kernel: LDS limit could not be raised");
    for (Factor* F : Fs) {
        if ((rc = rt_assemble(h, F))) return rc;
        F->nblocks = P.nblocks;
        const size_t bytes = (size_t)P.nblocks * n * n * elem;
        F->blocks = arena_take(h, bytes);
        if (!F->blocks)
            return fail(h, FDFD_ERR_OOM, "factor storage (" + std::to_string(bytes >> 20) +
                                             " MiB) does not fit in device memory");
        F->blocks_bytes = bytes;
        if (b3) {
            const size_t bb = (size_t)P.nblocks * b3_block_bytes(n);
            F->blocks_b = arena_take(h, bb);
            if (!F->blocks_b)
                return fail(h, FDFD_ERR_OOM, "split GEMM operand storage does not fit in device memory");
            F->blocks_b_bytes = bb;
        }
        F->triw = (double*)buf_alloc(h, sizeof(double) * P.tris.size() * (size_t)tri::work_doubles(n));
        F->prow = (int*)buf_alloc(h, sizeof(int) * (size_t)m * n);
        F->iprow = (int*)buf_alloc(h, sizeof(int) * (size_t)m * n);
        F->pivstat = (double*)buf_alloc(h, sizeof(double) * 3 * m);
        if (!F->triw || !F->prow || !F->iprow || !F->pivstat)
            return fail(h, FDFD_ERR_OOM, "pivot map storage does not fit in device memory");
    }
    // --- factorization job tables: per step, the jobs of factor 0, 1, ...
    std::vector<TriJob> tris;
    std::vector<SplitJob> splits;
    std::vector<int> nsplit;  // split jobs per GEMM step (all factors)
    std::vector<FormJob> forms;
    std::vector<LineJob> ljs;
    std::vector<GemmJob> gemms;
    // split-operand GEMM steps run as two launches: every job's first term,
    // then the second terms (accumulating), since terms carry their own scales
    std::vector<int> nsecond;
    for (const Step& st : P.fsteps) {
        std::vector<GemmJob> second;
        if (st.kind == ST_GEMM && b3) {
            // distinct operands of the step: A blocks as is, (B block, map) transposed
            int cnt = 0;
            for (int f = 0; f < B; ++f) {
                const Factor* F = Fs[f];
                std::vector<int> seenA;
                std::vector<std::pair<int, int>> seenB;
                for (int i = 0; i < st.count; ++i) {
                    const SGemm& g = P.gemms[st.first + i];
                    for (int t = 0; t < g.nt; ++t) {
                        if (std::find(seenA.begin(), seenA.end(), g.A[t]) == seenA.end()) {
                            seenA.push_back(g.A[t]);
                            splits.push_back({blk(F, g.A[t]), blk_b(F, g.A[t]), nullptr,
                                              blk_e(F, g.A[t]), 0});
                            ++cnt;
                        }
                        const std::pair<int, int> kb(g.B[t], g.bmap[t]);
                        if (std::find(seenB.begin(), seenB.end(), kb) == seenB.end()) {
                            seenB.push_back(kb);
                            splits.push_back({blk(F, g.B[t]), blk_b(F, g.B[t]),
                                              g.bmap[t] >= 0 ? F->prow + (long)g.bmap[t] * n : nullptr,
                                              blk_e(F, g.B[t]), 1});
                            ++cnt;
                        }
                    }
                }
            }
            nsplit.push_back(cnt);
        }
        for (int f = 0; f < B; ++f) {
            const Factor* F = Fs[f];
            auto dvec = [&](const SDiag& d, const double2** out2) {
                out2[0] = out2[1] = nullptr;
                for (int q = 0; q < 2; ++q)
                    if (d.line >= 0 && d.pl[q] >= 0) out2[q] = plane_ptr(F, N, d.line, d.pl[q]);
            };
            for (int i = 0; i < st.count; ++i) {
                if (st.kind == ST_TRI) {
                    const STri& s = P.tris[st.first + i];
                    TriJob J{};
                    J.out = blk(F, s.line);
                    own_planes(F, N, s.line, &J.sub, &J.diag, &J.sup);
                    dvec(s.dd, J.dd);
                    J.prow = F->prow + (long)s.line * n;
                    J.iprow = F->iprow + (long)s.line * n;
                    J.pivstat = F->pivstat + 3 * (long)s.line;
                    // kept per factor: the reduced solve applies T^-1 with it
                    J.work = F->triw + (size_t)(st.first + i) * tri::work_doubles(n);
                    J.si = F->si;
                    tris.push_back(J);
                } else if (st.kind == ST_FORM) {
                    const SForm& s = P.forms[st.first + i];
                    FormJob J{};
                    J.out = blk(F, s.out);
                    if (s.own >= 0) own_planes(F, N, s.own, &J.sub, &J.diag, &J.sup);
                    dvec(s.dd, J.dd);
                    J.nt = s.nt;
                    for (int t = 0; t < s.nt; ++t) {
                        J.Y[t] = blk(F, s.jl[t]);
                        J.prow[t] = F->prow + (long)s.jl[t] * n;
                        J.iprow[t] = F->iprow + (long)s.jl[t] * n;
                        dvec(s.rs[t], J.rs[t]);
                        dvec(s.cs[t], J.cs[t]);
                    }
                    J.si = F->si;
                    forms.push_back(J);
                } else if (st.kind == ST_GJ) {
                    const int j = P.gj[st.first + i];
                    const int c = f * st.count + i;  // work channel
                    LineJob J{};
                    J.Y = blk(F, j);
                    J.used = h->used + (long)c * n;
                    J.prow = F->prow + (long)j * n;
                    J.iprow = F->iprow + (long)j * n;
                    J.Vbuf = (char*)h->Vbuf + (size_t)c * n * VMAX * elem;
                    J.LU = (char*)h->LU + (size_t)c * 2 * PMAX * PMAX * elem;
                    J.piv = h->piv + c * VMAX;
                    J.Z = (char*)h->Zb + (size_t)c * VMAX * n * elem;
                    J.win = h->win + c * WIN;
                    J.verr = h->verr + f;  // the factor's own verification flag
                    J.pivstat = F->pivstat + 3 * (long)j;
                    J.n = n;
                    ljs.push_back(J);
                } else {
                    const SGemm& s = P.gemms[st.first + i];
                    GemmJob J{};
                    J.out = blk(F, s.out);
                    J.nt = s.nt;
                    J.acc = s.acc;
                    J.sub = s.sub;
                    if (b3) {  // split planes; the B gather is applied by the split
                        J.A[0] = blk_b(F, s.A[0]);
                        J.B[0] = blk_b(F, s.B[0]);
                        J.ea[0] = blk_e(F, s.A[0]);
                        J.eb[0] = blk_e(F, s.B[0]);
                        if (s.nt > 1) {
                            GemmJob J2 = J;
                            J2.nt = 0;  // counted with the first term
                            J2.acc = 1;
                            J2.A[0] = blk_b(F, s.A[1]);
                            J2.B[0] = blk_b(F, s.B[1]);
                            J2.ea[0] = blk_e(F, s.A[1]);
                            J2.eb[0] = blk_e(F, s.B[1]);
                            second.push_back(J2);
                        }
                    } else {
                        for (int t = 0; t < s.nt; ++t) {
                            J.A[t] = blk(F, s.A[t]);
                            J.B[t] = blk(F, s.B[t]);
                            J.bmap[t] = s.bmap[t] >= 0 ? F->prow + (long)s.bmap[t] * n : nullptr;
                        }
                    }
                    gemms.push_back(J);
                }
            }
        }
        if (st.kind == ST_GEMM && b3) {
            nsecond.push_back((int)second.size());
            gemms.insert(gemms.end(), second.begin(), second.end());
        }
    }
    Tables T(h);
    {
        TablePack pk;
        pk.add(&T.tri, tris);
        pk.add(&T.form, forms);
        pk.add(&T.lj, ljs);
        pk.add(&T.gemm, gemms);
        pk.add(&T.split, splits);
        if ((rc = pk.commit(h, &T.base))) return rc;
    }

    const bool fp64 = (prec == FDFD_FP64);
    const int NB = h->panel_width ? h->panel_width : PMAX;
    auto launch_all = [&](bool full, bool prof) {
        hipStream_t st = h->stream;
        auto scope = [&](int id, double fl, double by) {
            return prof ? new ProfScope(h, id, fl, by, st) : nullptr;
        };
        int ot = 0, of = 0, ol = 0, og = 0, os = 0, ogs = 0;
        for (const Step& s : P.fsteps) {
            const int nj = s.count * B;
            if (s.kind == ST_TRI) {
                ProfScope* ps = scope(P_SCHUR, 0, 3.0 * n * n * elem * nj);
                launch_trifactor(st, T.tri + ot, nj, n);
                if (fp64) launch_triinv<double>(st, T.tri + ot, nj, n);
                else launch_triinv<float>(st, T.tri + ot, nj, n);
                delete ps;
                ot += nj;
            } else if (s.kind == ST_FORM) {
                ProfScope* ps = scope(P_SCHUR, 0, (double)n * n * elem * nj);
                if (fp64) launch_form<double>(st, T.form + of, nj, n);
                else launch_form<float>(st, T.form + of, nj, n);
                delete ps;
                of += nj;
            } else if (s.kind == ST_GEMM) {
                if (b3) {
                    const int ns = nsplit[ogs++];
                    ProfScope* pss = scope(P_SCHUR, 0, 16.0 * n * n * ns);
                    launch_split(st, T.split + os, ns, n);
                    delete pss;
                    os += ns;
                }
                const double fl1 = 8.0 * n * (double)n * n;  // one complex n^3 product
                if (b3) {
                    // first terms, then the second terms accumulating into them
                    // (one profile scope per launch)
                    const int n2 = nsecond[ogs - 1];
                    ProfScope* ps = scope(P_GEMM, fl1 * nj, 0);
                    launch_gemm_b3(st, T.gemm + og, nj, n);
                    delete ps;
                    if (n2 > 0) {
                        ps = scope(P_GEMM, fl1 * n2, 0);
                        launch_gemm_b3(st, T.gemm + og + nj, n2, n);
                        delete ps;
                    }
                    og += nj + n2;
                } else {
                    double fl = 0;
                    for (int i = 0; i < nj; ++i) fl += fl1 * gemms[og + i].nt;
                    ProfScope* ps = scope(P_GEMM, fl, 0);
                    if (fp64) launch_gemm<double>(st, T.gemm + og, nj, n, h->gemm3m, h->gemm_waves);
                    else launch_gemm<float>(st, T.gemm + og, nj, n, h->gemm3m, h->gemm_waves);
                    delete ps;
                    og += nj;
                }
            } else {
                const LineJob* J = T.lj + ol;
                launch_gjreset(st, J, nj, n);
                // Panels of NB columns.  With the windowed search, two panels are
                // factored before one trailing pass applies both (the second panel's
                // columns, as the first panel leaves them, are formed by the window
                // panel for its window and by the Z build's check blocks for every
                // row, and its pivot rows by the Z build); the full search runs
                // single panels.
                for (int k0 = 0; k0 < n;) {
                    int nb1, nb2 = 0;
                    const int fresh = std::min(n, WIN + k0);
                    {
                        ProfScope* ps = scope(P_PANEL, 0, 0);
                        if (!full) {
                            nb1 = std::min(NB, n - k0);
                            if (fp64) launch_wpanel<double>(st, J, nj, n, k0, nb1, fresh, 0);
                            else launch_wpanel<float>(st, J, nj, n, k0, nb1, fresh, 0);
                        } else {
                            // full-search panel width: powers of two (16 columns of
                            // complex128 per thread would not fit 1024 threads' registers)
                            nb1 = fp64 ? 8 : 16;
                            while (nb1 > n - k0) nb1 >>= 1;
                            if (fp64) launch_panel<double>(st, J, nj, n, k0, nb1);
                            else launch_panel<float>(st, J, nj, n, k0, nb1);
                        }
                        delete ps;
                    }
                    {
                        ProfScope* ps = scope(P_ZBUILD, 0, 0);
                        if (fp64) launch_zbuild<double>(st, J, nj, n, k0, nb1, fresh, !full, 0, 0, 0);
                        else launch_zbuild<float>(st, J, nj, n, k0, nb1, fresh, !full, 0, 0, 0);
                        delete ps;
                    }
                    if (!full && nb1 == PMAX && k0 + nb1 < n && h->fuse_panels) {
                        const int k2 = k0 + nb1;
                        nb2 = std::min(NB, n - k2);
                        const int fresh2 = std::min(n, WIN + k2);
                        // (the second panel's values as the first panel leaves
                        // them are formed by the window panel, for its window,
                        // and by the Z build's check blocks, for every row)
                        ProfScope* ps = scope(P_PANEL, 0, 0);
                        if (fp64) launch_wpanel<double>(st, J, nj, n, k2, nb2, fresh2, 1);
                        else launch_wpanel<float>(st, J, nj, n, k2, nb2, fresh2, 1);
                        delete ps;
                        ps = scope(P_ZBUILD, 0, 0);
                        if (fp64) launch_zbuild<double>(st, J, nj, n, k2, nb2, fresh2, 1, 1, k0, nb1);
                        else launch_zbuild<float>(st, J, nj, n, k2, nb2, fresh2, 1, 1, k0, nb1);
                        delete ps;
                    }
                    {
                        ProfScope* ps = scope(P_TRAIL, 8.0 * n * (double)n * (nb1 + nb2) * nj,
                                              2.0 * n * (double)n * elem * nj);
                        if (fp64) launch_trailing<double>(st, J, nj, n, k0, nb1, nb2);
                        else if (h->trail_h3 && n % 4 == 0) launch_trailing_h3(st, J, nj, n, k0, nb1, nb2);
                        else launch_trailing<float>(st, J, nj, n, k0, nb1, nb2);
                        delete ps;
                    }
                    k0 += nb1 + nb2;
                }
                ol += nj;
            }
        }
    };

    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    // Pass 0: windowed pivot search with verification.  Pass 1 (only if a
    // row outside a window would have won a GEPP step, or a window pivot was
    // zero): full-column search, for the factors that failed.  When only some
    // of a batch failed, those are rebuilt on their own (fresh factors swapped
    // in), so the others keep their pass-0 result.
    int used_full = 0;
    std::vector<int> redo;           // factors rebuilt by the nested full-search build
    float ms = 0, ms_redo = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool full = (pass == 1) || h->full_pivot_search;
        used_full = full ? 1 : 0;
        HIPCHK(h, hipMemsetAsync(h->verr, 0, sizeof(int) * B, h->stream));
        HIPCHK(h, hipEventRecord(e0, h->stream));
        launch_all(full, h->prof.enabled);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventRecord(e1, h->stream));
        HIPCHK(h, hipEventSynchronize(e1));
        std::vector<int> verr(B, 0);
        COPY(h, verr.data(), h->verr, sizeof(int) * B, hipMemcpyDeviceToHost);
        if (pass == 0 && inject == FDFD_INJECT_PIVOT_FALLBACK) verr[0] = 1;
        std::vector<int> bad;
        for (int f = 0; f < B; ++f)
            if (verr[f]) bad.push_back(f);
        if (full || bad.empty()) break;
        if ((int)bad.size() < B) {
            HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
            std::vector<Factor*> sub;
            for (int f : bad) {
                Factor* G = new Factor();
                G->pol = Fs[f]->pol;
                G->omega = Fs[f]->omega;
                G->omega_pml = Fs[f]->omega_pml;
                G->shift = Fs[f]->shift;
                G->form = Fs[f]->form;
                G->precision = Fs[f]->precision;
                sub.push_back(G);
            }
            const bool was_full = h->full_pivot_search;
            h->full_pivot_search = true;
            rc = rt_build_factors(h, sub);
            h->full_pivot_search = was_full;
            if (rc) {
                for (Factor* G : sub) rt_free_factor(h, G);
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return rc;
            }
            ms_redo = sub[0]->t_factor_ms;
            for (size_t q = 0; q < bad.size(); ++q) {
                Factor* F = Fs[bad[q]];
                const int id = F->id;
                rt_release_factor_storage(h, F);
                *F = *sub[q];           // takes over the rebuilt factor's storage
                F->id = id;
                F->pivot_fallback = 1;
                delete sub[q];          // the struct only: its storage now belongs to F
                redo.push_back(bad[q]);
            }
            break;
        }
    }
    if (redo.empty()) HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    ms += ms_redo;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    auto rebuilt = [&](int f) { return std::find(redo.begin(), redo.end(), f) != redo.end(); };
    int first_singular = -1, singular_factor = -1, nsingular = 0;
    for (int fi = 0; fi < B; ++fi) {
        Factor* F = Fs[fi];
        F->t_factor_ms = ms;
        if (!rebuilt(fi)) F->pivot_fallback = used_full;
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
        const char* dbg = std::getenv("FDFD_DEBUG_PIVSTAT");
        if (dbg && (F->singular_line >= 0 || dbg[0] == '2')) {
            for (int j = 0; j < m; ++j)
                fprintf(stderr, "pivstat factor %d line %d: min %.3e max %.3e flag %g\n", fi, j,
                        ps[3 * j], ps[3 * j + 1], ps[3 * j + 2]);
        }
        if (inject == FDFD_INJECT_SINGULAR && F == Fs[0])
            F->singular_line = std::min(std::max(h->inject_arg, 0), m - 1);
        if (F->singular_line >= 0) ++nsingular;
        if (F->singular_line >= 0 && first_singular < 0) {
            first_singular = F->singular_line;
            singular_factor = fi;
        }
    }
    if (first_singular >= 0) {
        std::string msg = "zero pivot in line block " + std::to_string(first_singular);
        if (B > 1)
            msg += " of factor " + std::to_string(singular_factor) + " of " + std::to_string(B) +
                   " built together (omega " + std::to_string(Fs[singular_factor]->omega) +
                   "; " + std::to_string(nsingular) + " factors with a zero pivot)";
        if (used_full) msg += " after the full-column pivot search";
        return fail(h, FDFD_ERR_SINGULAR, msg);
    }

    // --- reduced-solve program of each factor (the rebuilt ones have theirs)
    for (int fi = 0; fi < B; ++fi) {
        if (rebuilt(fi)) continue;
        Factor* F = Fs[fi];
        std::vector<MvJob> mvs;
        std::vector<DiagJob> dgs;
        auto vec = [&](int v, int line) {
            double2* base = (v == VX) ? h->x_out : h->g_mid;
            return base + (long)line * F->sj;
        };
        for (const SMv& s : P.mvs) {
            MvJob J{};
            J.nt = s.nt;
            for (int t = 0; t < s.nt; ++t) {
                J.M[t] = blk(F, s.M[t]);
                J.imap[t] = s.imap[t] >= 0 ? F->prow + (long)s.imap[t] * n : nullptr;
                J.x[t] = vec(s.xv[t], s.xl[t]);
            }
            J.base = s.bl >= 0 ? vec(s.bv, s.bl) : nullptr;
            J.out = vec(s.ov, s.ol);
            J.omap = s.omap >= 0 ? F->iprow + (long)s.omap * n : nullptr;
            J.si = F->si;
            mvs.push_back(J);
        }
        for (const SDg& s : P.dgs) {
            DiagJob J{};
            J.out = vec(s.ov, s.ol);
            J.x0 = vec(s.x0v, s.x0l);
            J.nt = s.nt;
            for (int t = 0; t < s.nt; ++t) {
                J.x[t] = vec(s.xv[t], s.xl[t]);
                for (int q = 0; q < 2; ++q)
                    J.d[t][q] = (s.d[t].pl[q] >= 0) ? plane_ptr(F, N, s.d[t].line, s.d[t].pl[q])
                                                     : nullptr;
            }
            J.si = F->si;
            dgs.push_back(J);
        }
        std::vector<TrsJob> trs;
        for (const STrs& s : P.trss) {
            TrsJob J{};
            J.in = vec(s.xv, s.line);
            J.out = vec(s.ov, s.line);
            J.w = F->triw + (size_t)s.tri * tri::work_doubles(n);
            J.si = F->si;
            trs.push_back(J);
        }
        TablePack pk;
        pk.add(&F->mv_tab, mvs);
        pk.add(&F->dg_tab, dgs);
        pk.add(&F->trs_tab, trs);
        if ((rc = pk.commit(h, &F->tabs))) return rc;
        F->ssteps.clear();
        for (const Step& s : P.ssteps) {
            int mats = 0;
            if (s.kind == ST_MV)
                for (int i = 0; i < s.count; ++i) mats += P.mvs[s.first + i].nt;
            const int kind = s.kind == ST_MV ? 0 : (s.kind == ST_DIAG ? 1 : 2);
            F->ssteps.push_back({kind, s.first, s.count, mats});
        }
    }
    return FDFD_OK;
}

int rt_build_factor(Handle* h, Factor* F) { return rt_build_factors(h, std::vector<Factor*>{F}); }

// x_out <- A'^-1 x_out through the reduced-solve program
static void launch_solve(Handle* h, Factor* F, bool prof) {
    const int n = F->n;
    const bool fp64 = (F->precision == FDFD_FP64);
    for (const auto& s : F->ssteps) {
        ProfScope* ps = nullptr;
        if (prof) {
            // algorithmic bytes: the blocks streamed (matvec), the QR work and
            // two vectors (line solve), or the diagonal planes and vectors
            const double bytes = s.kind == 0   ? (double)n * n * F->elem * s.mats
                                 : s.kind == 2 ? (8.0 * tri::work_doubles(n) + 32.0 * n) * s.count
                                               : 64.0 * n * s.count;
            ps = new ProfScope(h, P_SWEEP, s.kind == 0 ? 8.0 * n * n * s.mats : 0, bytes);
        }
        if (s.kind == 0) {
            if (fp64) launch_mv<double>(h->stream, F->mv_tab + s.first, s.count, n);
            else launch_mv<float>(h->stream, F->mv_tab + s.first, s.count, n);
        } else if (s.kind == 2) {
            launch_trisolve(h->stream, F->trs_tab + s.first, s.count, n);
        } else {
            launch_diag(h->stream, F->dg_tab + s.first, s.count, n);
        }
        delete ps;
    }
}

// out = A'^-1 in, or A'^-T in = W^-1 A'^-1 W in (W below).  The
// launch sequence of a factor is captured once in a HIP graph and replayed
// (direct launches while per-kernel profiling is on).
int rt_prec_apply(Handle* h, Factor* F, int trans, const double2* in, double2* out) {
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    // A'^T = W^-1 A' W with W = sxf (x) syf (driven form) or sxb (x) syb
    // (sideband form): the periodic differences satisfy dxf^T = -dxb
    const double2* wx = F->form == 1 ? F->sxb : F->sxf;
    const double2* wy = F->form == 1 ? F->syb : F->syf;
    if (trans) launch_wscale(h->stream, h->x_out, in, wx, wy, g.Nx, g.Ny, 0);
    else
        HIPCHK(h, hipMemcpyAsync(h->x_out, in, sizeof(double2) * N, hipMemcpyDeviceToDevice,
                                 h->stream));
    if (h->prof.enabled || !h->use_graphs) {
        launch_solve(h, F, h->prof.enabled);
    } else {
        if (!F->solve_exec) {
            hipGraph_t gr;
            HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            launch_solve(h, F, false);
            HIPCHK(h, hipStreamEndCapture(h->stream, &gr));
            hipError_t e = hipGraphInstantiate(&F->solve_exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) {
                F->solve_exec = nullptr;
                return fail(h, FDFD_ERR_HIP, std::string("solve graph: ") + hipGetErrorString(e));
            }
        }
        HIPCHK(h, hipGraphLaunch(F->solve_exec, h->stream));
    }
    HIPCHK(h, hipGetLastError());
    if (trans) launch_wscale(h->stream, out, h->x_out, wx, wy, g.Nx, g.Ny, 1);
    else
        HIPCHK(h, hipMemcpyAsync(out, h->x_out, sizeof(double2) * N, hipMemcpyDeviceToDevice,
                                 h->stream));
    return FDFD_OK;
}

}  // namespace fdfd
