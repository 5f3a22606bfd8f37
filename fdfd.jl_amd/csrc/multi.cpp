// Lockstep flexible GMRES for several right-hand sides on one factor.
//
// The reference solves each source with its own solve() (src/solver/
// driven.jl:11-55; dolinearsolve, solver.jl:35).  Sources that share a
// factorization (one frequency, many sources) are solved together here: each
// FGMRES(30) iteration applies the preconditioner (the reduced solve of the
// block cyclic reduction) to every active source in one pass over the factor
// blocks, the operator to all of them in one stencil launch, and does the
// CGS2 orthogonalisation of all of them with one host round trip per
// reduction.  Every source keeps its own Krylov basis, Hessenberg matrix,
// restarts and convergence test; the arithmetic per source is that of the
// single-source FGMRES in runtime.cpp.  A source that converges (or
// stagnates) drops out of the launches.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "runtime.h"
#include "tri_qr.h"

namespace fdfd {

namespace {

constexpr int K = GMRES_KDIM;
constexpr int MNBLK = 256;  // reduction blocks per source
constexpr int PIN_SLOTS = 4;

// per-source device scratch: ridx | kcnt | coefficients | reduced values | partials
struct MSmall {
    int* ridx;
    int* kcnt;
    double2* coef;     // KMULTI x (K + 2)
    double2* outc;     // KMULTI x (K + 2)
    double* outr;      // KMULTI
    double2* part;     // KMULTI x MNBLK x (K + 2)
    double* partr;     // KMULTI x MNBLK
};
size_t msmall_bytes() {
    return sizeof(int) * 2 * KMULTI + sizeof(double2) * 2 * KMULTI * (K + 2) +
           sizeof(double) * KMULTI + sizeof(double2) * (size_t)KMULTI * MNBLK * (K + 2) +
           sizeof(double) * KMULTI * MNBLK + 1024;
}
MSmall msmall(void* base) {
    char* p = (char*)base;
    MSmall m;
    m.ridx = (int*)p; p += sizeof(int) * KMULTI;
    m.kcnt = (int*)p; p += sizeof(int) * KMULTI;
    p = (char*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    m.coef = (double2*)p; p += sizeof(double2) * KMULTI * (K + 2);
    m.outc = (double2*)p; p += sizeof(double2) * KMULTI * (K + 2);
    m.outr = (double*)p; p += sizeof(double) * KMULTI;
    m.part = (double2*)p; p += sizeof(double2) * (size_t)KMULTI * MNBLK * (K + 2);
    m.partr = (double*)p;
    return m;
}
// pinned host staging: PIN_SLOTS slots of (2 KMULTI ints + KMULTI (K + 2) complex)
struct PinSlot {
    int* ints;
    double2* c;
};
size_t pin_slot_bytes() { return sizeof(int) * 2 * KMULTI + sizeof(double2) * KMULTI * (K + 2) + 64; }
PinSlot pin_slot(void* base, int i) {
    char* p = (char*)base + (size_t)i * pin_slot_bytes();
    PinSlot s;
    s.ints = (int*)p;
    s.c = (double2*)(((uintptr_t)(p + sizeof(int) * 2 * KMULTI) + 15) & ~(uintptr_t)15);
    return s;
}

}  // namespace

int ensure_multi_work(Handle* h, long N, int k) {
    if (!h->msmall) {
        HIPCHK(h, hipMalloc(&h->msmall, msmall_bytes()));
        HIPCHK(h, hipHostMalloc(&h->mpin, pin_slot_bytes() * PIN_SLOTS, hipHostMallocDefault));
    }
    if (h->multi_N >= N && h->multi_k >= k) return FDFD_OK;
    N = std::max(N, h->multi_N);
    k = std::max(k, h->multi_k);
    for (double2** p : {&h->mV, &h->mZ, &h->mB, &h->mX}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    h->multi_N = 0;
    h->multi_k = 0;
    if (hipMalloc(&h->mV, sizeof(double2) * (size_t)(K + 1) * k * N) != hipSuccess ||
        hipMalloc(&h->mZ, sizeof(double2) * (size_t)K * k * N) != hipSuccess ||
        hipMalloc(&h->mB, sizeof(double2) * (size_t)k * N) != hipSuccess ||
        hipMalloc(&h->mX, sizeof(double2) * (size_t)k * N) != hipSuccess) {
        (void)hipGetLastError();
        for (double2** p : {&h->mV, &h->mZ, &h->mB, &h->mX}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        return fail(h, FDFD_ERR_OOM, "multi-source Krylov bases do not fit in device memory");
    }
    h->multi_N = N;
    h->multi_k = k;
    return FDFD_OK;
}

void free_multi_work(Handle* h) {
    for (double2** p : {&h->mV, &h->mZ, &h->mB, &h->mX}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (h->msmall) (void)hipFree(h->msmall);
    if (h->mpin) (void)hipHostFree(h->mpin);
    if (h->mpack) (void)hipFree(h->mpack);
    h->msmall = h->mpin = h->mpack = nullptr;
    h->mpack_bytes = 0;
    h->multi_N = 0;
    h->multi_k = 0;
}

// x_out slots ridx <- A'^-1 x_out slots ridx, through the factor's reduced-solve program
static void launch_solve_multi(Handle* h, Factor* F, long rs, const int* ridx, int na, bool prof) {
    const int n = F->n;
    const bool fp64 = (F->precision == FDFD_FP64);
    for (const auto& s : F->ssteps) {
        ProfScope* ps = nullptr;
        if (prof) {
            const double bytes = s.kind == 0   ? (double)n * n * F->elem * s.mats
                                 : s.kind == 2 ? (8.0 * tri::work_doubles(n) + 32.0 * n * na) * s.count
                                               : 64.0 * n * s.count * na;
            ps = new ProfScope(h, P_SWEEP, s.kind == 0 ? 8.0 * n * n * s.mats * na : 0, bytes);
        }
        if (s.kind == 0) {
            const bool mf = h->mv_mfma != 0;
            if (fp64)
                launch_mv_multi<double>(h->stream, F->mv_tab + s.first, s.count, n, rs, ridx, na, mf,
                                        h->mpack);
            else
                launch_mv_multi<float>(h->stream, F->mv_tab + s.first, s.count, n, rs, ridx, na, mf,
                                       h->mpack);
        } else if (s.kind == 2) {
            launch_trisolve_multi(h->stream, F->trs_tab + s.first, s.count, n, rs, ridx, na);
        } else {
            launch_diag_multi(h->stream, F->dg_tab + s.first, s.count, n, rs, ridx, na);
        }
        delete ps;
    }
}

int rt_gmres_multi(Handle* h, Factor* F, int nr, const double2* B, double2* X, double tol,
                   int* iters, double* worst) {
    if (nr < 1 || nr > KMULTI) return fail(h, FDFD_ERR_INVALID, "1 <= sources <= 16 per multi solve");
    if (F->form != 0 && F->form != 1)
        return fail(h, FDFD_ERR_INVALID, "multi-source solves take driven or sideband factors");
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    int rc;
    if ((rc = ensure_multi_work(h, N, nr))) return rc;
    // packed-input panel of the widest matrix-vector step (MFMA path)
    int maxjobs = 0;
    for (const auto& s : F->ssteps)
        if (s.kind == 0) maxjobs = std::max(maxjobs, s.count);
    const size_t pack = mv_pack_bytes(maxjobs, F->n, F->precision == FDFD_FP64);
    if (h->mv_mfma && pack > h->mpack_bytes) {
        if (h->mpack) (void)hipFree(h->mpack);
        h->mpack = nullptr;
        h->mpack_bytes = 0;
        if (hipMalloc(&h->mpack, pack) != hipSuccess) {
            (void)hipGetLastError();
            return fail(h, FDFD_ERR_OOM, "multi-source packed inputs do not fit in device memory");
        }
        h->mpack_bytes = pack;
    }
    if (mv_multi_prepare<float>(F->n) || mv_multi_prepare<double>(F->n) ||
        trisolve_multi_prepare(F->n))
        return fail(h, FDFD_ERR_HIP, "multi-source solve kernels: LDS limit could not be raised");
    const MSmall m = msmall(h->msmall);
    const long ldv = (long)nr * N;  // basis vector j of source r at j * ldv + r * N
    double2* V = h->mV;
    double2* Z = h->mZ;
    hipStream_t st = h->stream;
    const StencilOp op{F->pol, F->form, g.Nx, g.Ny, h->eps, F->sxf, F->sxb, F->syf, F->syb,
                       (g.dx * g.dx) / (g.dy * g.dy), F->kdiag};
    const bool prof = h->prof.enabled;
    int slot = 0;
    auto next_slot = [&]() { PinSlot s = pin_slot(h->mpin, slot); slot = (slot + 1) % PIN_SLOTS; return s; };
    // upload the active list (and per-source counts) for the following launches
    std::vector<int> act;
    auto upload_list = [&](const std::vector<int>& lst, const std::vector<int>* cnt) -> int {
        PinSlot s = next_slot();
        for (size_t i = 0; i < lst.size(); ++i) s.ints[i] = lst[i];
        if (cnt) for (size_t i = 0; i < lst.size(); ++i) s.ints[KMULTI + i] = (*cnt)[i];
        HIPCHK(h, hipMemcpyAsync(m.ridx, s.ints, sizeof(int) * lst.size(), hipMemcpyHostToDevice, st));
        if (cnt)
            HIPCHK(h, hipMemcpyAsync(m.kcnt, s.ints + KMULTI, sizeof(int) * lst.size(),
                                     hipMemcpyHostToDevice, st));
        return FDFD_OK;
    };
    auto norms = [&](const double2* x, long rs, int na, std::vector<double>& out) -> int {
        launch_norm2_b(st, x, rs, N, m.partr, m.outr, MNBLK, m.ridx, na);
        out.resize(na);
        HIPCHK(h, hipMemcpyAsync(out.data(), m.outr, sizeof(double) * na, hipMemcpyDeviceToHost, st));
        HIPCHK(h, hipStreamSynchronize(st));
        for (auto& v : out) v = std::sqrt(v);
        return FDFD_OK;
    };
    // scale source slots: y_r = a_r x_r
    auto scale = [&](double2* y, const double2* x, const std::vector<double2>& a) -> int {
        PinSlot s = next_slot();
        for (size_t i = 0; i < a.size(); ++i) s.c[i] = a[i];
        HIPCHK(h, hipMemcpyAsync(m.coef, s.c, sizeof(double2) * a.size(), hipMemcpyHostToDevice, st));
        launch_scale_b(st, y, N, x, N, m.coef, N, m.ridx, (int)a.size());
        return FDFD_OK;
    };
    auto apply = [&](double2* y, const double2* x, const double2* b, int na) {
        ProfScope* ps = prof ? new ProfScope(h, P_STENCIL, 0, (b ? 64.0 : 48.0) * N * na) : nullptr;
        launch_stencil(st, y, x, op, 0, b, N, m.ridx, na);
        delete ps;
    };

    std::vector<double> bnorm(nr), resid(nr, 1.0), prev(nr, 2.0);
    std::vector<int> state(nr, 0);  // 0 iterating, 1 converged, 2 failed
    act.resize(nr);
    for (int r = 0; r < nr; ++r) act[r] = r;
    if ((rc = upload_list(act, nullptr))) return rc;
    {
        std::vector<double> nb;
        if ((rc = norms(B, N, nr, nb))) return rc;
        for (int r = 0; r < nr; ++r) bnorm[r] = nb[r];
    }
    for (int r = 0; r < nr; ++r)
        if (bnorm[r] == 0) {
            launch_zero(st, X + (long)r * N, N);
            state[r] = 1;
            resid[r] = 0;
        }
    struct Src {
        std::vector<std::complex<double>> H, gv, sn, y;
        std::vector<double> cs;
        int jdone = 0;
        bool ext = false;
    };
    std::vector<Src> S(nr);
    for (auto& s : S) {
        s.H.assign((size_t)(K + 1) * K, 0);
        s.gv.assign(K + 1, 0);
        s.sn.assign(K, 0);
        s.y.assign(K, 0);
        s.cs.assign(K, 0);
    }
    const int max_restarts = 40;
    for (int rs = 0; rs <= max_restarts; ++rs) {
        act.clear();
        for (int r = 0; r < nr; ++r)
            if (state[r] == 0) act.push_back(r);
        if (act.empty()) break;
        if ((rc = upload_list(act, nullptr))) return rc;
        const int na = (int)act.size();
        // V_0,r = B_r - A' X_r, beta_r = ||V_0,r||
        apply(V, X, B, na);
        std::vector<double> beta;
        if ((rc = norms(V, N, na, beta))) return rc;
        std::vector<int> cyc;
        std::vector<double2> inv;
        for (int q = 0; q < na; ++q) {
            const int r = act[q];
            resid[r] = beta[q] / bnorm[r];
            if (resid[r] <= tol) { state[r] = 1; continue; }
            // no restart left, or a whole cycle that gained less than 10%: the
            // rounding floor of the operator
            if (rs == max_restarts || (rs > 0 && resid[r] > 0.9 * prev[r])) { state[r] = 2; continue; }
            prev[r] = resid[r];
            cyc.push_back(r);
            inv.push_back(make_double2(1.0 / beta[q], 0.0));
            Src& s = S[r];
            std::fill(s.gv.begin(), s.gv.end(), 0.0);
            s.gv[0] = beta[q];
            s.jdone = 0;
            s.ext = true;
        }
        if (cyc.empty()) continue;
        if ((rc = upload_list(cyc, nullptr))) return rc;
        if ((rc = scale(V, V, inv))) return rc;
        std::vector<int> it = cyc;
        for (int j = 0; j < K && !it.empty(); ++j) {
            const int ni = (int)it.size();
            if ((rc = upload_list(it, nullptr))) return rc;
            double2* Vj = V + (long)j * ldv;
            double2* Zj = Z + (long)j * ldv;
            double2* W = V + (long)(j + 1) * ldv;
            // Z_j = M^-1 V_j (through the x_out slots of the reduced solve)
            std::vector<double2> one(ni, make_double2(1.0, 0.0));
            if ((rc = scale(h->x_out, Vj, one))) return rc;
            launch_solve_multi(h, F, N, m.ridx, ni, prof);
            if ((rc = scale(Zj, h->x_out, one))) return rc;
            // W = A' Z_j, then CGS2 against V_0..V_j
            apply(W, Zj, nullptr, ni);
            std::vector<std::vector<std::complex<double>>> hsum(ni, std::vector<std::complex<double>>(j + 1, 0));
            for (int pass = 0; pass < 2; ++pass) {
                launch_multidot_b(st, V, ldv, N, j + 1, W, N, N, m.part, m.outc, MNBLK, m.ridx, ni);
                PinSlot s = next_slot();
                HIPCHK(h, hipMemcpyAsync(s.c, m.outc, sizeof(double2) * ni * (j + 1),
                                         hipMemcpyDeviceToHost, st));
                HIPCHK(h, hipStreamSynchronize(st));
                std::vector<int> cnt(ni, j + 1);
                PinSlot u = next_slot();
                for (int q = 0; q < ni; ++q)
                    for (int i = 0; i <= j; ++i) {
                        hsum[q][i] += std::complex<double>(s.c[q * (j + 1) + i].x, s.c[q * (j + 1) + i].y);
                        u.c[q * (K + 2) + i] = s.c[q * (j + 1) + i];
                    }
                for (int q = 0; q < ni; ++q) u.ints[KMULTI + q] = j + 1;
                HIPCHK(h, hipMemcpyAsync(m.kcnt, u.ints + KMULTI, sizeof(int) * ni, hipMemcpyHostToDevice, st));
                HIPCHK(h, hipMemcpyAsync(m.coef, u.c, sizeof(double2) * ni * (K + 2), hipMemcpyHostToDevice, st));
                launch_gemv_update_b(st, W, N, V, ldv, N, m.kcnt, m.coef, K + 2, -1.0, N, m.ridx, ni);
            }
            std::vector<double> hn;
            if ((rc = norms(W, N, ni, hn))) return rc;
            std::vector<double2> hinv(ni);
            for (int q = 0; q < ni; ++q) hinv[q] = make_double2(hn[q] > 0 ? 1.0 / hn[q] : 1.0, 0.0);
            if ((rc = scale(W, W, hinv))) return rc;
            std::vector<int> still;
            for (int q = 0; q < ni; ++q) {
                const int r = it[q];
                Src& s = S[r];
                auto Hij = [&](int i, int jj) -> std::complex<double>& { return s.H[(size_t)i * K + jj]; };
                for (int i = 0; i <= j; ++i) Hij(i, j) = hsum[q][i];
                Hij(j + 1, j) = hn[q];
                for (int i = 0; i < j; ++i) {
                    const std::complex<double> a = Hij(i, j), bb = Hij(i + 1, j);
                    Hij(i, j) = s.cs[i] * a + s.sn[i] * bb;
                    Hij(i + 1, j) = -std::conj(s.sn[i]) * a + s.cs[i] * bb;
                }
                const std::complex<double> a = Hij(j, j);
                const double bb = hn[q];
                const double t = std::sqrt(std::norm(a) + bb * bb);
                if (std::abs(a) == 0) {
                    s.cs[j] = 0;
                    s.sn[j] = 1;
                    Hij(j, j) = bb;
                } else {
                    const std::complex<double> ph = a / std::abs(a);
                    s.cs[j] = std::abs(a) / t;
                    s.sn[j] = ph * bb / t;
                    Hij(j, j) = ph * t;
                }
                s.gv[j + 1] = -std::conj(s.sn[j]) * s.gv[j];
                s.gv[j] = s.cs[j] * s.gv[j];
                ++*iters;
                s.jdone = j + 1;
                const double est = std::abs(s.gv[j + 1]) / bnorm[r];
                if (!(est <= 0.5 * tol || hn[q] == 0)) still.push_back(r);
            }
            it = still;
        }
        // y = H^-1 g and X_r += Z_r y_r for every source of the cycle
        PinSlot u = next_slot();
        for (size_t q = 0; q < cyc.size(); ++q) {
            Src& s = S[cyc[q]];
            for (int i = s.jdone - 1; i >= 0; --i) {
                std::complex<double> v = s.gv[i];
                for (int l = i + 1; l < s.jdone; ++l) v -= s.H[(size_t)i * K + l] * s.y[l];
                s.y[i] = v / s.H[(size_t)i * K + i];
            }
            for (int i = 0; i < s.jdone; ++i) u.c[q * (K + 2) + i] = make_double2(s.y[i].real(), s.y[i].imag());
            u.ints[q] = cyc[q];
            u.ints[KMULTI + q] = s.jdone;
        }
        HIPCHK(h, hipMemcpyAsync(m.ridx, u.ints, sizeof(int) * cyc.size(), hipMemcpyHostToDevice, st));
        HIPCHK(h, hipMemcpyAsync(m.kcnt, u.ints + KMULTI, sizeof(int) * cyc.size(), hipMemcpyHostToDevice, st));
        HIPCHK(h, hipMemcpyAsync(m.coef, u.c, sizeof(double2) * cyc.size() * (K + 2), hipMemcpyHostToDevice, st));
        launch_gemv_update_b(st, X, N, Z, ldv, N, m.kcnt, m.coef, K + 2, 1.0, N, m.ridx, (int)cyc.size());
        HIPCHK(h, hipStreamSynchronize(st));  // the pinned slots are reused next cycle
    }
    double w = 0;
    int failed = 0;
    for (int r = 0; r < nr; ++r) {
        w = std::max(w, resid[r]);
        if (state[r] != 1) ++failed;
    }
    *worst = std::max(*worst, w);
    HIPCHK(h, hipGetLastError());
    if (failed) {
        char buf[200];
        snprintf(buf, sizeof(buf),
                 "refinement did not reach the tolerance %.1e for %d of %d sources (worst relative "
                 "residual %.3e)", tol, failed, nr, w);
        return fail(h, FDFD_ERR_NOCONV, buf);
    }
    return FDFD_OK;
}

}  // namespace fdfd
