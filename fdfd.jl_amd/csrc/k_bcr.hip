// Block cyclic reduction kernels of libfdfd (gfx950).
//
// The lines of the grid form a ring (the y-periodic corner blocks couple the
// first and last line).  Each reduction level eliminates every other line of
// the current ring: the eliminated line's block is inverted (Gauss-Jordan,
// k_factor.hip) and its couplings fold into its two surviving neighbours,
//   D_s -= C_sj X_j C_js,   C_new(s, u) = -C_sj X_j C_ju.
// On the first level the couplings are diagonal and the folds are
// elementwise (k_form); afterwards they are dense and run as batched complex
// GEMMs on MFMA (k_gemm) in the order Z = C_sj X_j, then D_s -= Z C_js.
// The reduced solve applies the same steps to a vector (k_mv, k_diag).
#include "fdfd_common.h"

#include <atomic>
#include "k_device.h"
#include "tri_qr.h"

namespace fdfd {

// ---------------------------------------------------------------------------
// First-level block formation: one workgroup per output row a.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_form(const FormJob* jobs, int n) {
    using C = cplx<R>;
    const FormJob J = jobs[blockIdx.y];
    __shared__ C row0[MAX_LINE];
    __shared__ C row1[MAX_LINE];
    const int a = blockIdx.x;
    const long si = J.si;
    if (J.nt > 0) {
        const C* y = (const C*)J.Y[0] + (long)J.prow[0][a] * n;
        for (int c = threadIdx.x; c < n; c += blockDim.x) row0[c] = y[c];
    }
    if (J.nt > 1) {
        const C* y = (const C*)J.Y[1] + (long)J.prow[1][a] * n;
        for (int c = threadIdx.x; c < n; c += blockDim.x) row1[c] = y[c];
    }
    __syncthreads();
    const long ea = (long)a * si;
    const double2 z = make_double2(0, 0);
    double2 dg = J.diag ? J.diag[ea] : z;
    const double2 up = J.sup ? J.sup[ea] : z, dn = J.sub ? J.sub[ea] : z;
    if (J.dd[0]) dg = cadd(dg, J.dd[0][ea]);
    if (J.dd[1]) dg = cadd(dg, J.dd[1][ea]);
    double2 r0 = z, r1 = z;
    if (J.nt > 0) {
        r0 = J.rs[0][0][ea];
        if (J.rs[0][1]) r0 = cadd(r0, J.rs[0][1][ea]);
    }
    if (J.nt > 1) {
        r1 = J.rs[1][0][ea];
        if (J.rs[1][1]) r1 = cadd(r1, J.rs[1][1][ea]);
    }
    const int ap = (a + 1 == n) ? 0 : a + 1, am = (a == 0) ? n - 1 : a - 1;
    C* out = (C*)J.out + (long)a * n;
    for (int b = threadIdx.x; b < n; b += blockDim.x) {
        const long eb = (long)b * si;
        double2 v = z;
        if (b == a) v = cadd(v, dg);
        if (b == ap) v = cadd(v, up);
        if (b == am) v = cadd(v, dn);
        if (J.nt > 0) {
            double2 c0 = J.cs[0][0][eb];
            if (J.cs[0][1]) c0 = cadd(c0, J.cs[0][1][eb]);
            v = csub(v, cmul(cmul(r0, c0), to_d(row0[J.iprow[0][b]])));
        }
        if (J.nt > 1) {
            double2 c1 = J.cs[1][0][eb];
            if (J.cs[1][1]) c1 = cadd(c1, J.cs[1][1][eb]);
            v = csub(v, cmul(cmul(r1, c1), to_d(row1[J.iprow[1][b]])));
        }
        out[b] = from_d<C>(v);
    }
}

template <typename R>
void launch_form(hipStream_t s, const FormJob* jobs_d, int njobs, int n) {
    k_form<R><<<dim3(n, njobs), 256, 0, s>>>(jobs_d, n);
}
template void launch_form<double>(hipStream_t, const FormJob*, int, int);
template void launch_form<float>(hipStream_t, const FormJob*, int, int);

// ---------------------------------------------------------------------------
// First-level inversion of the lines' own periodic tridiagonal blocks
// (Givens QR, tri_qr.h): one thread factors one line (O(n), FP64), then one
// thread per column forms X[:, b] = R^-1 Q e_b (O(n) each), writing the
// column twice (Q e_b, then the solution) in coalesced rows of a wave.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_trifactor(const TriJob* jobs, int njobs, int n) {
    const int jid = blockIdx.x * 64 + threadIdx.x;
    if (jid >= njobs) return;
    const TriJob J = jobs[jid];
    auto get = [&](int i, tri::cz& d, tri::cz& e, tri::cz& f) {
        const long ei = (long)i * J.si;
        double2 dv = J.diag[ei];
        if (J.dd[0]) dv = cadd(dv, J.dd[0][ei]);
        if (J.dd[1]) dv = cadd(dv, J.dd[1][ei]);
        const double2 ev = J.sup[ei], fv = J.sub[ei];
        d = tri::mkz(dv.x, dv.y);
        e = tri::mkz(ev.x, ev.y);
        f = tri::mkz(fv.x, fv.y);
    };
    double pmin, pmax;
    const int rc = tri::factor(n, get, J.work, pmin, pmax);
    J.pivstat[0] = pmin;
    J.pivstat[1] = pmax;
    J.pivstat[2] = rc ? 1.0 : 0.0;
}

void launch_trifactor(hipStream_t s, const TriJob* jobs_d, int njobs, int n) {
    k_trifactor<<<(njobs + 63) / 64, 64, 0, s>>>(jobs_d, njobs, n);
}

constexpr int TRI_U = 8;  // back-substitution steps whose operands are loaded together

template <typename R>
__global__ __launch_bounds__(64) void k_triinv(const TriJob* jobs, int n) {
    using C = cplx<R>;
    const TriJob J = jobs[blockIdx.y];
    const int b0 = blockIdx.x * 64, b = b0 + threadIdx.x;
    // wave-uniform start of the rotation pass (the column n-1 is dense)
    const int kstart = (b0 + 63 >= n - 1) ? 0 : tri::first_nonzero(n, b0);
    if (b >= n) return;
    C* X = (C*)J.out;
    const double* w = J.work;
    auto put = [&](int k, tri::cz v) { gstore(X + (long)k * n + b, mk<C>(R(v.re), R(v.im))); };
    auto get = [&](int k) {
        const C y = gload(X + (long)k * n + b);
        return tri::mkz((double)y.x, (double)y.y);
    };
    tri::apply_q(n, w, b, kstart, put);
    // x = R^-1 y: the arithmetic of tri::apply_rinv, with the y[k] of TRI_U
    // steps loaded before the steps run (y[k] is overwritten only at step k)
    // and the rows of R of those steps, the same for every column of the
    // line, staged in LDS by the wave (one wave per workgroup), so the back
    // substitution does not wait on each load
    using tri::cz;
    __shared__ double rsh[TRI_U * tri::RROW];
    const int lane = threadIdx.x;  // b0 + lane = b < n
    const int k0 = tri::first_nonzero(n, b);
    auto y = [&](int k, cz v) { return k >= k0 ? v : tri::mkz(0, 0); };
    const cz xB = tri::zmul(y(n - 1, get(n - 1)), tri::get_r(w, n, n - 1, 0));
    put(n - 1, xB);
    if (n > 1) {
        const cz xA = tri::zmul(tri::zsub(y(n - 2, get(n - 2)), tri::zmul(tri::get_r(w, n, n - 2, 4), xB)),
                                tri::get_r(w, n, n - 2, 0));
        put(n - 2, xA);
        cz x1 = xA, x2 = xB;  // x[k+1], x[k+2]
        const double* wr = w + tri::ROT * n;
        // lanes past n have returned: the nact active lanes stage the rows
        const int nact = min(64, n - b0);
        for (int kt = n - 3; kt >= 0; kt -= TRI_U) {
            cz yy[TRI_U];
#pragma unroll
            for (int u = 0; u < TRI_U; ++u) yy[u] = get(max(kt - u, 0));
            for (int e = lane; e < TRI_U * tri::RROW; e += nact)
                rsh[e] = gload_s(wr + (long)tri::RROW * max(kt - e / tri::RROW, 0) + e % tri::RROW);
            __syncthreads();
#pragma unroll
            for (int u = 0; u < TRI_U; ++u) {
                const int k = kt - u;
                if (k >= 0) {
                    const double* r = rsh + tri::RROW * u;
                    cz v = y(k, yy[u]);
                    v = tri::zsub(v, tri::zmul(tri::mkz(r[2], r[3]), x1));
                    v = tri::zsub(v, tri::zmul(tri::mkz(r[4], r[5]), x2));
                    v = tri::zsub(v, tri::zmul(tri::mkz(r[6], r[7]), xA));
                    v = tri::zsub(v, tri::zmul(tri::mkz(r[8], r[9]), xB));
                    const cz x = tri::zmul(v, tri::mkz(r[0], r[1]));
                    put(k, x);
                    x2 = x1;
                    x1 = x;
                }
            }
            __syncthreads();  // the rows are read before the next batch is staged
        }
    }
    J.prow[b] = b;
    J.iprow[b] = b;
}

template <typename R>
void launch_triinv(hipStream_t s, const TriJob* jobs_d, int njobs, int n) {
    k_triinv<R><<<dim3((n + 63) / 64, njobs), 64, 0, s>>>(jobs_d, n);
}
template void launch_triinv<double>(hipStream_t, const TriJob*, int, int);
template void launch_triinv<float>(hipStream_t, const TriJob*, int, int);

// First-level line solve of the reduced solve, one line per workgroup of two
// waves: out = T^-1 in with T's Givens QR (tri_qr.h, FP64): y = Q in, then
// out = R^-1 y, both in place on the line's vector in LDS.  The two passes
// are sequential recurrences over the cells, run by lane 0 of wave 0 with
// the same arithmetic, in the same order, as tri::apply_q_vec and
// tri::apply_rinv; wave 1 meanwhile stages the next TS_CH steps of the
// factor (rotations, then rows of R, which are contiguous in the work array)
// into the other half of a double buffer, so each step reads its operands
// from LDS and the memory latency is hidden behind the recurrence.  Lane 0
// loads the LDS operands of TS_U steps at a time before running them, so one
// LDS round trip serves a batch instead of every step; the steps of a batch
// are straight-line code (the last two cells' special steps are peeled off),
// so the independent parts of neighbouring steps interleave.
constexpr int TS_CH = 128;     // steps per staged chunk
constexpr int TS_STEP = 10;    // doubles per step: 6 (rotations) or 10 (row of R)
constexpr int TS_U = 8;        // steps whose LDS operands are loaded together

__global__ __launch_bounds__(128) void k_trisolve(const TrsJob* jobs, int n) {
    extern __shared__ double tsm[];  // vector (2 n doubles), then 2 x TS_CH x TS_STEP
    double2* v = (double2*)tsm;
    double* buf = tsm + 2 * n;
    const TrsJob J = jobs[blockIdx.x];
    const int t = threadIdx.x;
    const bool comp = t == 0, loader = t >= 64;
    const int ll = t - 64;
    // stage doubles [src0, src0 + cnt) of the work array into buffer half b
    // (cnt <= TS_CH * TS_STEP): all loads in flight before the LDS stores
    auto stage = [&](int b, long src0, int cnt) {
        constexpr int PER = TS_CH * TS_STEP / 64;
        double r[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 64 + ll;
            r[q] = gload_s(J.w + src0 + (e < cnt ? e : 0));
        }
        double* d = buf + b * TS_CH * TS_STEP;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 64 + ll;
            if (e < cnt) d[e] = r[q];
        }
    };
    for (int k0 = 0; k0 < n; k0 += 4 * 128) {
        double2 r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + q * 128 + t;
            r[q] = gload(J.in + (k < n ? (long)k : 0) * J.si);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + q * 128 + t;
            if (k < n) v[k] = r[q];
        }
    }
    using tri::cz;
    auto ld = [&](int k) { return tri::mkz(v[k].x, v[k].y); };
    auto st = [&](int k, cz z) { v[k] = make_double2(z.re, z.im); };
    // pass 1, y = Q in: rotation steps k = 0 .. n-2 (6 doubles each at w[6 k])
    const int n1 = n - 1;
    const int nc1 = (n1 + TS_CH - 1) / TS_CH;
    if (loader && n1 > 0) stage(0, 0, tri::ROT * min(TS_CH, n1));
    __syncthreads();
    cz carry = tri::mkz(0, 0), tail = tri::mkz(0, 0);
    if (comp && n > 1) {
        carry = ld(0);
        tail = ld(n - 1);
    }
    for (int c = 0; c < nc1; ++c) {
        const int k0 = c * TS_CH, len = min(TS_CH, n1 - k0);
        if (loader && c + 1 < nc1) {
            const int k1 = k0 + TS_CH;
            stage((c + 1) & 1, (long)tri::ROT * k1, tri::ROT * min(TS_CH, n1 - k1));
        }
        if (comp) {
            const double* rb = buf + (c & 1) * TS_CH * TS_STEP;
            // general steps k < n - 2 (rot1 on rows (k, k+1), rot2 on rows
            // (k, n-1)); the final rotation (k = n - 2) is peeled off
            const int lenG = max(0, min(len, n - 2 - k0));
            auto gstep = [&](const double* r, cz nxt, int k) {
                tri::rot(r[0], tri::mkz(r[1], r[2]), carry, nxt);
                tri::rot(r[3], tri::mkz(r[4], r[5]), carry, tail);
                st(k, carry);
                carry = nxt;
            };
            // full batches of TS_U steps as straight-line code (no branch
            // between steps, so the independent parts of neighbouring steps
            // interleave): every operand of a batch (rotations and v[k + 1];
            // v[k + 1] is stored at step k + 1, after it is read) is loaded
            // before the first step
            int i0 = 0;
            for (; i0 + TS_U <= lenG; i0 += TS_U) {
                double rr[TS_U][tri::ROT];
                cz nx[TS_U];
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
#pragma unroll
                    for (int j = 0; j < tri::ROT; ++j) rr[u][j] = rb[tri::ROT * (i0 + u) + j];
                    nx[u] = ld(k0 + i0 + u + 1);
                }
#pragma unroll
                for (int u = 0; u < TS_U; ++u) gstep(rr[u], nx[u], k0 + i0 + u);
            }
            for (; i0 < lenG; ++i0) gstep(rb + tri::ROT * i0, ld(k0 + i0 + 1), k0 + i0);
            if (k0 + len - 1 == n - 2) {  // k = n-2: rot1 on rows (n-2, n-1)
                const double* r = rb + tri::ROT * (len - 1);
                tri::rot(r[0], tri::mkz(r[1], r[2]), carry, tail);
                st(n - 2, carry);
                st(n - 1, tail);
            }
        }
        __syncthreads();
    }
    // pass 2, out = R^-1 y, rows k = n-1 .. 0 (10 doubles each at w[6 n + 10 k]:
    // 1/R[k][k], R[k][k+1], R[k][k+2], R[k][n-2], R[k][n-1])
    const long rbase = (long)tri::ROT * n;
    const int nc2 = (n + TS_CH - 1) / TS_CH;
    if (loader) {
        const int lo = max(0, n - TS_CH);
        stage(0, rbase + (long)tri::RROW * lo, tri::RROW * (n - lo));
    }
    __syncthreads();
    cz x1 = tri::mkz(0, 0), x2 = tri::mkz(0, 0), xA = tri::mkz(0, 0), xB = tri::mkz(0, 0);
    for (int c = 0; c < nc2; ++c) {
        const int hi = n - c * TS_CH, lo = max(0, hi - TS_CH);
        if (loader && c + 1 < nc2) {
            const int lo2 = max(0, lo - TS_CH);
            stage((c + 1) & 1, rbase + (long)tri::RROW * lo2, tri::RROW * (lo - lo2));
        }
        if (comp) {
            const double* rb = buf + (c & 1) * TS_CH * TS_STEP;
            auto row = [&](int k) { return rb + tri::RROW * (k - lo); };
            int kt = hi - 1;
            if (c == 0) {  // rows n-1 and n-2 give xB and xA
                xB = tri::zmul(ld(n - 1), tri::mkz(row(n - 1)[0], row(n - 1)[1]));
                st(n - 1, xB);
                if (n >= 2) {
                    const double* r = row(n - 2);
                    xA = tri::zmul(tri::zsub(ld(n - 2), tri::zmul(tri::mkz(r[8], r[9]), xB)),
                                   tri::mkz(r[0], r[1]));
                    st(n - 2, xA);
                    x1 = xA;
                    x2 = xB;
                }
                kt = n - 3;
            }
            auto gstep = [&](const double* r, cz yk, int k) {
                cz y = yk;
                y = tri::zsub(y, tri::zmul(tri::mkz(r[2], r[3]), x1));
                y = tri::zsub(y, tri::zmul(tri::mkz(r[4], r[5]), x2));
                y = tri::zsub(y, tri::zmul(tri::mkz(r[6], r[7]), xA));
                y = tri::zsub(y, tri::zmul(tri::mkz(r[8], r[9]), xB));
                const cz x = tri::zmul(y, tri::mkz(r[0], r[1]));
                st(k, x);
                x2 = x1;
                x1 = x;
            };
            // full batches of TS_U rows going down as straight-line code:
            // rows of R and y[k] (stored only at step k) loaded first
            for (; kt - (TS_U - 1) >= lo; kt -= TS_U) {
                double rr[TS_U][tri::RROW];
                cz yy[TS_U];
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
#pragma unroll
                    for (int j = 0; j < tri::RROW; ++j) rr[u][j] = row(kt - u)[j];
                    yy[u] = ld(kt - u);
                }
#pragma unroll
                for (int u = 0; u < TS_U; ++u) gstep(rr[u], yy[u], kt - u);
            }
            for (; kt >= lo; --kt) gstep(row(kt), ld(kt), kt);
        }
        __syncthreads();
    }
    for (int k = t; k < n; k += 128) gstore(J.out + (long)k * J.si, v[k]);
}

static size_t trisolve_lds(int n) { return sizeof(double) * (2 * (size_t)n + 2 * TS_CH * TS_STEP); }

int trisolve_prepare(int n) {
    // set once per device for the largest line length seen (a launch for a
    // shorter line uses less)
    static std::atomic<int> done[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = -1;
    if (dev >= 0 && done[dev].load() >= n) return (int)hipSuccess;
    const hipError_t e = hipFuncSetAttribute((const void*)k_trisolve,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)trisolve_lds(n));
    if (e == hipSuccess && dev >= 0) {
        int cur = done[dev].load();
        while (cur < n && !done[dev].compare_exchange_weak(cur, n)) {
        }
    }
    return (int)e;
}

void launch_trisolve(hipStream_t s, const TrsJob* jobs_d, int njobs, int n) {
    k_trisolve<<<njobs, 128, trisolve_lds(n), s>>>(jobs_d, n);
}

// ---------------------------------------------------------------------------
// Gauss-Jordan start state of a line: no row pivoted, window = rows 0..63.
// ---------------------------------------------------------------------------
__global__ void k_gjreset(const LineJob* jobs, int n) {
    const LineJob J = jobs[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) {
        J.used[r] = 0;
        J.iprow[r] = -1;
    }
    if (r < WIN) J.win[r] = (r < n) ? r : -1;
    if (r == 0) {
        J.pivstat[0] = __builtin_huge_val();
        J.pivstat[1] = 0;
        J.pivstat[2] = 0;
    }
}

void launch_gjreset(hipStream_t s, const LineJob* jobs_d, int njobs, int n) {
    const int t = n > WIN ? n : WIN;
    k_gjreset<<<dim3((t + 255) / 256, njobs), 256, 0, s>>>(jobs_d, n);
}

// ---------------------------------------------------------------------------
// Batched complex GEMM on MFMA:  out = [acc ? out : 0] -/+ sum_t A_t B_t[bmap_t]
// Workgroup tile TM x TN (2 x 2 waves), K staged GKT at a time through LDS
// (double-buffered, re and im planes split, A stored k-major so that both
// operands are read as MFMA fragments without bank conflicts).  The complex
// product is four real MFMAs per 16x16 block:
//   re += Ar Br - Ai Bi,  im += Ar Bi + Ai Br,
// or, with G3 (Gauss's three-multiplication form), three:
//   t1 += Ar Br, t2 += Ai Bi, t3 += (Ar + Ai)(Br + Bi);
//   re = t1 - t2, im = t3 - t1 - t2.
// ---------------------------------------------------------------------------
template <typename R> struct gemm_cfg;
template <> struct gemm_cfg<float> {
    static constexpr int TM = 128, TN = 128;
    static constexpr int PA = TM + 16, ROT = 0;
};
// complex128: the A loader's four threads of a row write k-rows 4 apart in the
// same cycle; rotating each group of four k-rows by 8 columns (ROT) spreads
// them over the banks, and a row pitch of TM + 48 doubles keeps the fragment
// reads of neighbouring k-rows in different bank halves
template <> struct gemm_cfg<double> {
    static constexpr int TM = 128, TN = 64;
    static constexpr int PA = TM + 48, ROT = 8;
};
constexpr int GKT = 16;

// WR x WC waves per workgroup, each owning a (TM/WR) x (TN/WC) sub-tile.
template <typename R, bool G3, int WR, int WC>
__global__ __launch_bounds__(64 * WR * WC) void k_gemm(const GemmJob* jobs, int n) {
    using C = cplx<R>;
    using M = mfma_t<R>;
    using acc_t = typename M::acc;
    constexpr int NT = 64 * WR * WC;
    constexpr int TM = gemm_cfg<R>::TM, TN = gemm_cfg<R>::TN;
    constexpr int WM = TM / WR, WN = TN / WC;
    constexpr int BM = WM / 16, BN = WN / 16;
    constexpr int PA = gemm_cfg<R>::PA, PB = TN + 16, ROT = gemm_cfg<R>::ROT;
    // column of A row i at k-row k of a stage
    auto acol = [](int k, int i) { return i + ROT * ((k >> 2) & 3); };
    constexpr int AE = TM * GKT / NT;   // A elements per thread per stage (consecutive k)
    constexpr int TPR = GKT / AE;       // threads per A row
    constexpr int TPK = NT / GKT;       // threads per B k-row
    constexpr int BE = TN / TPK;        // B elements per thread per stage (stride TPK)
    static_assert(AE * TPR == GKT && BE * TPK == TN && BM >= 1 && BN >= 1, "tile shape");
    __shared__ R sAr[2][GKT][PA], sAi[2][GKT][PA];
    __shared__ R sBr[2][GKT][PB], sBi[2][GKT][PB];
    const GemmJob J = jobs[blockIdx.z];
    const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = wv / WC, wc = wv % WC;
    const int lr = lane & 15, lk = lane >> 4;
    // loader coordinates
    const int arow = t / TPR, ak = (t % TPR) * AE;  // A: row, first of AE k's
    const int bk = t / TPK, bc = t % TPK;           // B: k row, column lane (stride TPK)
    const int spt = (n + GKT - 1) / GKT;
    const int ns = J.nt * spt;
    C ra[AE], rb[BE];
    auto load = [&](int s) {
        const int ts = s / spt, kb = (s - ts * spt) * GKT;
        const C* A = (const C*)(ts ? J.A[1] : J.A[0]);
        const C* B = (const C*)(ts ? J.B[1] : J.B[0]);
        const int* bm = ts ? J.bmap[1] : J.bmap[0];
        const int i = i0 + arow;
        // (guarded assignments: a select between two complex structs would
        // be lowered through a stack slot)
        if (i < n && kb + ak + AE <= n) {
            const C* src = A + (long)i * n + kb + ak;
#pragma unroll
            for (int e = 0; e < AE; ++e) ra[e] = gload(src + e);
        } else {
#pragma unroll
            for (int e = 0; e < AE; ++e) {
                const int k = kb + ak + e;
                ra[e].x = R(0);
                ra[e].y = R(0);
                if (i < n && k < n) ra[e] = gload(A + (long)i * n + k);
            }
        }
        const int k = kb + bk;
        const C* srcb = B + (long)((k < n) ? (bm ? gload_s(bm + k) : k) : 0) * n;
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            const int j = j0 + bc + TPK * e;
            rb[e].x = R(0);
            rb[e].y = R(0);
            if (k < n && j < n) rb[e] = gload(srcb + j);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < AE; ++e) {
            sAr[buf][ak + e][acol(ak + e, arow)] = ra[e].x;
            sAi[buf][ak + e][acol(ak + e, arow)] = ra[e].y;
        }
#pragma unroll
        for (int e = 0; e < BE; ++e) {
            sBr[buf][bk][bc + TPK * e] = rb[e].x;
            sBi[buf][bk][bc + TPK * e] = rb[e].y;
        }
    };
    // G3: cre = t1, cim = t3, c2 = t2; otherwise cre = re, cim = im
    acc_t cre[BM][BN], cim[BM][BN], c2[G3 ? BM : 1][G3 ? BN : 1];
#pragma unroll
    for (int bi = 0; bi < BM; ++bi)
#pragma unroll
        for (int bj = 0; bj < BN; ++bj) {
            cre[bi][bj] = acc_t{0, 0, 0, 0};
            cim[bi][bj] = acc_t{0, 0, 0, 0};
            if constexpr (G3) c2[bi][bj] = acc_t{0, 0, 0, 0};
        }
    // the products of one stage (operands in LDS buffer cur)
    auto compute = [&](int cur) {
#pragma unroll
        for (int kk = 0; kk < GKT; kk += 4) {
            R ar[BM], an[BM], ai[BM], br[BN], bi[BN];
#pragma unroll
            for (int q = 0; q < BM; ++q) {
                ar[q] = sAr[cur][kk + lk][acol(kk + lk, wr * WM + q * 16 + lr)];
                ai[q] = sAi[cur][kk + lk][acol(kk + lk, wr * WM + q * 16 + lr)];
                an[q] = -ai[q];
            }
#pragma unroll
            for (int q = 0; q < BN; ++q) {
                br[q] = sBr[cur][kk + lk][wc * WN + q * 16 + lr];
                bi[q] = sBi[cur][kk + lk][wc * WN + q * 16 + lr];
            }
            // passes over all blocks so that consecutive MFMAs use different
            // accumulators
            if constexpr (G3) {
                R as[BM], bs[BN];
#pragma unroll
                for (int p = 0; p < BM; ++p) as[p] = ar[p] + ai[p];
#pragma unroll
                for (int q = 0; q < BN; ++q) bs[q] = br[q] + bi[q];
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) cre[p][q] = M::run(ar[p], br[q], cre[p][q]);
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) c2[p][q] = M::run(ai[p], bi[q], c2[p][q]);
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) cim[p][q] = M::run(as[p], bs[q], cim[p][q]);
            } else {
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) cre[p][q] = M::run(ar[p], br[q], cre[p][q]);
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) cim[p][q] = M::run(ar[p], bi[q], cim[p][q]);
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) cre[p][q] = M::run(an[p], bi[q], cre[p][q]);
#pragma unroll
                for (int p = 0; p < BM; ++p)
#pragma unroll
                    for (int q = 0; q < BN; ++q) cim[p][q] = M::run(ai[p], br[q], cim[p][q]);
            }
        }
    };
    // Ping-pong stages (as in k_gemm_b3): between two barriers a wave
    // multiplies the current stage and stores the next one (loaded into
    // registers one interval earlier) into the other buffer, then issues the
    // loads of the stage after it; the two waves of each SIMD (wv and wv + 4
    // with 8 waves) do these in opposite orders, so one's memory work
    // overlaps the other's MFMAs.  Every wave passes the same barriers.
    load(0);
    store(0);
    if (ns > 1) load(1);
    __syncthreads();
    const bool mem_first = WR * WC == 8 && wv >= 4;
    for (int s = 0; s < ns; ++s) {
        const int cur = s & 1;
        if (mem_first) {
            if (s + 1 < ns) {
                store(cur ^ 1);
                if (s + 2 < ns) load(s + 2);
            }
            compute(cur);
        } else {
            compute(cur);
            if (s + 1 < ns) {
                store(cur ^ 1);
                if (s + 2 < ns) load(s + 2);
            }
        }
        __syncthreads();
    }
    C* out = (C*)J.out;
#pragma unroll
    for (int p = 0; p < BM; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + wr * WM + p * 16 + M::row(lane, v);
            if (i >= n) continue;
#pragma unroll
            for (int q = 0; q < BN; ++q) {
                const int j = j0 + wc * WN + q * 16 + lr;
                if (j >= n) continue;
                C* o = out + (long)i * n + j;
                R ox = R(0), oy = R(0);
                if (J.acc) {
                    const C old = gload(o);
                    ox = old.x;
                    oy = old.y;
                }
                R re, im;
                if constexpr (G3) {
                    re = cre[p][q][v] - c2[p][q][v];
                    im = cim[p][q][v] - cre[p][q][v] - c2[p][q][v];
                } else {
                    re = cre[p][q][v];
                    im = cim[p][q][v];
                }
                gstore(o, J.sub ? mk<C>(ox - re, oy - im) : mk<C>(ox + re, oy + im));
            }
        }
}

template <typename R>
void launch_gemm(hipStream_t s, const GemmJob* jobs_d, int njobs, int n, int gauss3, int waves) {
    constexpr int TM = gemm_cfg<R>::TM, TN = gemm_cfg<R>::TN;
    const dim3 grid((n + TN - 1) / TN, (n + TM - 1) / TM, njobs);
    if (waves == 8) {
        if (gauss3) k_gemm<R, true, 4, 2><<<grid, 512, 0, s>>>(jobs_d, n);
        else k_gemm<R, false, 4, 2><<<grid, 512, 0, s>>>(jobs_d, n);
    } else {
        if (gauss3) k_gemm<R, true, 2, 2><<<grid, 256, 0, s>>>(jobs_d, n);
        else k_gemm<R, false, 2, 2><<<grid, 256, 0, s>>>(jobs_d, n);
    }
}
template void launch_gemm<double>(hipStream_t, const GemmJob*, int, int, int, int);
template void launch_gemm<float>(hipStream_t, const GemmJob*, int, int, int, int);

// ---------------------------------------------------------------------------
// Three-product GEMMs for the mixed path (its default).  Every FP32 operand
// row is scaled by a power of two 2^e (so its largest |re|, |im| is at most
// 2^14) and split once into fp16 hi = f16(2^e x), lo = f16(2^e x - hi); a real
// product is hi hi + hi lo + lo hi accumulated in FP32 by
// v_mfma_f32_16x16x32_f16 (about 22 bits of each operand: relative error
// ~2^-21 per product, close to an FP32 GEMM), and each accumulated term is
// unscaled exactly by 2^-(e_row + e_col).  The row scaling keeps fp16's range
// (2^-24 .. 65504) around each row's largest entry.
// ---------------------------------------------------------------------------
// One workgroup per 16-row strip (tile row block rb) of an operand: as is,
// strip rows are block rows 16 rb + r and k runs over the columns;
// transposed, strip rows are block columns c = 16 rb + r and k runs over the
// block rows through the gather, (r, k) = src[map[k]][c].  The strip is read
// once into registers in the tile layout -- wave wv holds tiles kb = wv,
// wv + SP_W, .. (at most SP_TPW for n <= MAX_LINE), lane l row l % 16 and
// k = 32 kb + 8 (l / 16) + [0, 8) -- then each row's largest |re|, |im|
// (lanes l, l ^ 16, l ^ 32, l ^ 48 of all SP_W waves) gives its exponent and
// the scaled fp16 planes are written, one wave per 1 KB tile.
constexpr int SP_W = 8;  // waves per strip
constexpr int SP_TPW = (MAX_LINE / 32 + SP_W - 1) / SP_W;

__global__ __launch_bounds__(64 * SP_W) void k_split(const SplitJob* jobs, int n) {
    const SplitJob J = jobs[blockIdx.y];
    const int rb = blockIdx.x, r0 = 16 * rb;
    const int nkb = b3_nkb(n);
    const float2* src = (const float2*)J.src;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int r = lane & 15, row = r0 + r, kq = 8 * (lane >> 4);
    __shared__ float wmax[SP_W][16];
    float2 v[SP_TPW][8];
    // loads: clamped addresses, no branches (entries past n are masked below)
#pragma unroll
    for (int q = 0; q < SP_TPW; ++q) {
        const int kb = wv + SP_W * q;
        if (kb < nkb) {
            if (!J.trans) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = 32 * kb + kq + e;
                    v[q][e] = gload(src + ((row < n && k < n) ? (long)row * n + k : 0));
                }
            } else {
                int mr[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = 32 * kb + kq + e;
                    mr[e] = (J.map && k < n) ? gload_s(J.map + k) : k;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = 32 * kb + kq + e;
                    v[q][e] = gload(src + ((row < n && k < n) ? (long)mr[e] * n + row : 0));
                }
            }
        }
    }
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < SP_TPW; ++q) {
        const int kb = wv + SP_W * q;
        if (kb < nkb) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = 32 * kb + kq + e;
                const bool ok = row < n && k < n;
                v[q][e].x = ok ? v[q][e].x : 0.f;
                v[q][e].y = ok ? v[q][e].y : 0.f;
                m = fmaxf(m, fmaxf(fabsf(v[q][e].x), fabsf(v[q][e].y)));
            }
        }
    }
    m = fmaxf(m, __shfl_xor(m, 16, 64));
    m = fmaxf(m, __shfl_xor(m, 32, 64));
    if (lane < 16) wmax[wv][lane] = m;
    __syncthreads();
    float mx = wmax[0][r];
#pragma unroll
    for (int w = 1; w < SP_W; ++w) mx = fmaxf(mx, wmax[w][r]);
    const int ex = split_exp(mx);
    if (t < 16) J.exps[r0 + t] = ex;
    const float sc = exp2i(ex);
    const long plane = b3_plane(n);
    _Float16* dst = (_Float16*)J.dst;
#pragma unroll
    for (int q = 0; q < SP_TPW; ++q) {
        const int kb = wv + SP_W * q;
        if (kb < nkb) {
            h8v rh, rl, ih, il;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                _Float16 h, l;
                split_f16(v[q][e].x * sc, h, l);
                rh[e] = h;
                rl[e] = l;
                split_f16(v[q][e].y * sc, h, l);
                ih[e] = h;
                il[e] = l;
            }
            const long o = ((long)rb * nkb + kb) * 512 + 8 * lane;
            *(GLOBAL_AS h8v*)(dst + o) = rh;
            *(GLOBAL_AS h8v*)(dst + plane + o) = rl;
            *(GLOBAL_AS h8v*)(dst + 2 * plane + o) = ih;
            *(GLOBAL_AS h8v*)(dst + 3 * plane + o) = il;
        }
    }
}

void launch_split(hipStream_t s, const SplitJob* jobs_d, int njobs, int n) {
    k_split<<<dim3(b3_nrb(n), njobs), 64 * SP_W, 0, s>>>(jobs_d, n);
}

// GEMM on split operands (tiled planes, b3_plane): A as [row][k], B^T as
// [col][k] with the gather already applied.  Complex products in the
// four-multiplication form with Ai Bi kept in its own accumulator:
//   re = Ar Br - Ai Bi,  im = Ar Bi + Ai Br,
// each real product as hi hi + hi lo + lo hi on v_mfma_f32_16x16x32_f16.
// Tile 128 x 128, 2 x 4 waves of 64 x 32, K staged 32 at a time through
// double-buffered LDS.  Every operand access of a stage (global load, LDS
// store, fragment read) is one wave reading or writing a contiguous 1 KB
// tile, 16 bytes per lane, which keeps the LDS accesses conflict-free.  The
// flat grid is ordered so each XCD (workgroups are dealt to the 8 XCDs
// round-robin) works through a contiguous range of jobs and keeps their
// operand panels in its own L2.  A job is one term out = [acc ? out : 0] -/+
// A B: the terms of a sum carry their own row scales, so the host issues a
// sum of two as two launches, the second accumulating into the first.
constexpr int B3_TM = 128, B3_TN = 128;
constexpr int EP_BLK = 4;  // row blocks whose old values the epilogue loads together (1, 2: same or slower)

__global__ __launch_bounds__(512) void k_gemm_b3(const GemmJob* jobs, int n, int ntx, int total) {
    __shared__ __align__(16) h8v sA[2][4][8][64];  // [buffer][plane][row block][lane]
    __shared__ __align__(16) h8v sB[2][4][8][64];
    const int L = blockIdx.x, per = total / 8;
    const int lin = (L < 8 * per) ? (L % 8) * per + L / 8 : L;
    const int job = lin / (ntx * ntx), rem = lin % (ntx * ntx);
    const GemmJob J = jobs[job];
    const int i0 = (rem / ntx) * B3_TM, j0 = (rem % ntx) * B3_TN;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = wv >> 2, wc = wv & 3;
    const int lr = lane & 15, lg = lane >> 4;
    const int nkb = b3_nkb(n), nrb = b3_nrb(n);
    const long plane = b3_plane(n);
    // loader: per stage and operand 4 planes x 8 row blocks of 1 KB, four
    // 16-byte pieces per thread (a wave loads one whole tile)
    h8v ra[4], rb[4];
    auto load = [&](int kb) {
        const _Float16* A = (const _Float16*)J.A[0];
        const _Float16* B = (const _Float16*)J.B[0];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int id = t + 512 * e;
            const int pl = id >> 9, blk = (id >> 6) & 7, l = id & 63;
            const int ta = i0 / 16 + blk, tb = j0 / 16 + blk;
            // past the last row block the loads read tile 0 (zeroed at the store)
            ra[e] = *(const GLOBAL_AS h8v*)(A + (ta < nrb ? pl * plane + ((long)ta * nkb + kb) * 512 + 8 * l : 0));
            rb[e] = *(const GLOBAL_AS h8v*)(B + (tb < nrb ? pl * plane + ((long)tb * nkb + kb) * 512 + 8 * l : 0));
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int id = t + 512 * e;
            const int pl = id >> 9, blk = (id >> 6) & 7, l = id & 63;
            sA[buf][pl][blk][l] = (i0 / 16 + blk < nrb) ? ra[e] : h8v{};
            sB[buf][pl][blk][l] = (j0 / 16 + blk < nrb) ? rb[e] : h8v{};
        }
    };
    f4v cre[4][2], cim[4][2], cii[4][2];
    auto zero = [&]() {
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                cre[p][q] = f4v{0, 0, 0, 0};
                cim[p][q] = f4v{0, 0, 0, 0};
                cii[p][q] = f4v{0, 0, 0, 0};
            }
    };
    float2* out = (float2*)J.out;
    // add the term (unscaled) to the output tile: the 16 row exponents are
    // loaded first, then one row block (4 rows) at a time, its old values
    // all in flight before its stores (a scheduling barrier after each block
    // keeps the old values of one block, beside the live accumulators, per
    // lane)
    auto epilogue = [&]() {
        const int* ea = J.ea[0];
        const int* eb = J.eb[0];
        const bool rd = J.acc;
        const int jb = j0 + wc * 32 + lr;
        int ej[2], ei[4][4];
#pragma unroll
        for (int q = 0; q < 2; ++q) ej[q] = gload_s(eb + (jb + 16 * q < n ? jb + 16 * q : 0));
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + wr * 64 + p * 16 + 4 * lg + v;
                ei[p][v] = gload_s(ea + (i < n ? i : 0));
            }
#pragma unroll
        for (int pp = 0; pp < 4; pp += EP_BLK) {
            float2 old[EP_BLK][4][2];
            if (rd) {  // uniform; clamped addresses, no select after a load
#pragma unroll
                for (int pb = 0; pb < EP_BLK; ++pb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = i0 + wr * 64 + (pp + pb) * 16 + 4 * lg + v;
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            const bool ok = i < n && jb + 16 * q < n;
                            old[pb][v][q] = gload(out + (ok ? (long)i * n + jb + 16 * q : 0));
                        }
                    }
            } else {
#pragma unroll
                for (int pb = 0; pb < EP_BLK; ++pb)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
#pragma unroll
                        for (int q = 0; q < 2; ++q) old[pb][v][q] = make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int pb = 0; pb < EP_BLK; ++pb) {
                const int p = pp + pb;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = i0 + wr * 64 + p * 16 + 4 * lg + v;
                    if (i < n) {
                        float2* orow = out + (long)i * n + jb;
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            if (jb + 16 * q < n) {
                                const float sc = exp2i(-(ei[p][v] + ej[q]));
                                const float re = (cre[p][q][v] - cii[p][q][v]) * sc, im = cim[p][q][v] * sc;
                                const float2 o = old[pb][v][q];
                                gstore(orow + 16 * q, J.sub ? make_float2(o.x - re, o.y - im) : make_float2(o.x + re, o.y + im));
                            }
                        }
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    auto compute = [&](int cur) {
        h8v a[4][4], b[2][4];  // [block][plane]: re hi, re lo, im hi, im lo
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int v = 0; v < 4; ++v) a[p][v] = sA[cur][v][wr * 4 + p][lane];
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) b[q][v] = sB[cur][v][wc * 2 + q][lane];
        // (x_hi y_hi + x_hi y_lo + x_lo y_hi) for the four real products,
        // terms outermost: each term round runs the four products over all
        // 8 blocks, so an accumulator's next MFMA is 16-32 MFMAs away
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int xa = (r == 2) ? 1 : 0, yb = (r == 1) ? 1 : 0;  // hi hi, hi lo, lo hi
#define B3_MM(ACC, XP, YP)                                                                      \
    _Pragma("unroll") for (int p = 0; p < 4; ++p) _Pragma("unroll") for (int q = 0; q < 2; ++q) \
        ACC[p][q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[p][XP + xa], b[q][YP + yb], ACC[p][q], 0, 0, 0);
            B3_MM(cre, 0, 0)  // Ar Br
            B3_MM(cim, 0, 2)  // Ar Bi
            B3_MM(cii, 2, 2)  // Ai Bi
            B3_MM(cim, 2, 0)  // Ai Br
        }
#undef B3_MM
    };
    // Ping-pong stages.  Between two barriers a wave multiplies the current
    // stage from one LDS buffer and stores the next stage (loaded into
    // registers one interval earlier) into the other, then issues the global
    // loads of the stage after it.  The two waves of a SIMD belong to the two
    // row halves (wr 0 and 1) and do these in opposite orders: wr = 1 stores
    // and loads first, so its memory work overlaps the MFMAs of its wr = 0
    // partner, whose memory work then overlaps wr = 1's MFMAs.  Every wave
    // passes the same barriers.
    zero();
    load(0);
    store(0);
    if (nkb > 1) load(1);
    __syncthreads();
    const bool mem_first = wr == 1;
    for (int s = 0; s < nkb; ++s) {
        const int cur = s & 1;
        if (mem_first) {
            if (s + 1 < nkb) {
                store(cur ^ 1);
                if (s + 2 < nkb) load(s + 2);
            }
            compute(cur);
        } else {
            compute(cur);
            if (s + 1 < nkb) {
                store(cur ^ 1);
                if (s + 2 < nkb) load(s + 2);
            }
        }
        __syncthreads();
    }
    epilogue();
}

void launch_gemm_b3(hipStream_t s, const GemmJob* jobs_d, int njobs, int n) {
    const int ntx = (n + B3_TM - 1) / B3_TM;
    const int total = ntx * ntx * njobs;
    k_gemm_b3<<<total, 512, 0, s>>>(jobs_d, n, ntx, total);
}

// ---------------------------------------------------------------------------
// Batched dense matrix-vector step of the reduced solve.  The gathered input
// vectors sit in LDS; each wave computes two rows at a time, four columns
// per lane per trip (eight loads in flight per lane).
// ---------------------------------------------------------------------------
template <typename R>
__device__ inline void mv_rows(const cplx<R>* ya, const cplx<R>* yb, const cplx<R>* w, int n,
                               int lane, R& ar, R& ai, R& br, R& bi) {
    using C = cplx<R>;
    int c = lane;
    for (; c + 192 < n; c += 256) {
        const C a0 = gload(ya + c), a1 = gload(ya + c + 64), a2 = gload(ya + c + 128),
                a3 = gload(ya + c + 192);
        const C b0 = gload(yb + c), b1 = gload(yb + c + 64), b2 = gload(yb + c + 128),
                b3 = gload(yb + c + 192);
        const C x0 = w[c], x1 = w[c + 64], x2 = w[c + 128], x3 = w[c + 192];
        ar += a0.x * x0.x - a0.y * x0.y + a1.x * x1.x - a1.y * x1.y + a2.x * x2.x -
              a2.y * x2.y + a3.x * x3.x - a3.y * x3.y;
        ai += a0.x * x0.y + a0.y * x0.x + a1.x * x1.y + a1.y * x1.x + a2.x * x2.y + a2.y * x2.x +
              a3.x * x3.y + a3.y * x3.x;
        br += b0.x * x0.x - b0.y * x0.y + b1.x * x1.x - b1.y * x1.y + b2.x * x2.x -
              b2.y * x2.y + b3.x * x3.x - b3.y * x3.y;
        bi += b0.x * x0.y + b0.y * x0.x + b1.x * x1.y + b1.y * x1.x + b2.x * x2.y + b2.y * x2.x +
              b3.x * x3.y + b3.y * x3.x;
    }
    for (; c < n; c += 64) {
        const C a = gload(ya + c), b = gload(yb + c), x = w[c];
        ar += a.x * x.x - a.y * x.y;
        ai += a.x * x.y + a.y * x.x;
        br += b.x * x.x - b.y * x.y;
        bi += b.x * x.y + b.y * x.x;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_mv(const MvJob* jobs, int n) {
    using C = cplx<R>;
    const MvJob J = jobs[blockIdx.y];
    __shared__ C w0[MAX_LINE];
    __shared__ C w1[MAX_LINE];
    const long si = J.si;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int s0 = J.imap[0] ? J.imap[0][c] : c;
        w0[c] = from_d<C>(J.x[0][(long)s0 * si]);
        if (J.nt > 1) {
            const int s1 = J.imap[1] ? J.imap[1][c] : c;
            w1[c] = from_d<C>(J.x[1][(long)s1 * si]);
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rows_per_block = (n + gridDim.x - 1) / gridDim.x;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    for (int r = r0 + 2 * wv; r < r1; r += 8) {
        const bool two = (r + 1) < r1;
        const int rb = two ? r + 1 : r;
        R ar = 0, ai = 0, br = 0, bi = 0;
        {
            const C* Mt = (const C*)J.M[0];
            mv_rows<R>(Mt + (long)r * n, Mt + (long)rb * n, w0, n, lane, ar, ai, br, bi);
        }
        if (J.nt > 1) {
            const C* Mt = (const C*)J.M[1];
            mv_rows<R>(Mt + (long)r * n, Mt + (long)rb * n, w1, n, lane, ar, ai, br, bi);
        }
        for (int o = 32; o > 0; o >>= 1) {
            ar += shfl_x(ar, o);
            ai += shfl_x(ai, o);
            br += shfl_x(br, o);
            bi += shfl_x(bi, o);
        }
        if (lane == 0) {
            long e = (long)(J.omap ? J.omap[r] : r) * si;
            double2 z = make_double2((double)ar, (double)ai);
            J.out[e] = J.base ? csub(J.base[e], z) : z;
            if (two) {
                e = (long)(J.omap ? J.omap[r + 1] : r + 1) * si;
                z = make_double2((double)br, (double)bi);
                J.out[e] = J.base ? csub(J.base[e], z) : z;
            }
        }
    }
}

template <typename R>
void launch_mv(hipStream_t s, const MvJob* jobs_d, int njobs, int n) {
    // enough workgroups to cover the chip, at least 8 rows per workgroup
    int nblk = (4096 + njobs - 1) / njobs;
    nblk = std::max(1, std::min(nblk, (n + 7) / 8));
    k_mv<R><<<dim3(nblk, njobs), 256, 0, s>>>(jobs_d, n);
}
template void launch_mv<double>(hipStream_t, const MvJob*, int, int);
template void launch_mv<float>(hipStream_t, const MvJob*, int, int);

// ---------------------------------------------------------------------------
// First-level (diagonal) step of the reduced solve.
// ---------------------------------------------------------------------------
__global__ void k_diag(const DiagJob* jobs, int n) {
    const DiagJob J = jobs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long e = (long)i * J.si;
    double2 v = J.x0[e];
    if (J.nt > 0) {
        double2 d = J.d[0][0][e];
        if (J.d[0][1]) d = cadd(d, J.d[0][1][e]);
        v = csub(v, cmul(d, J.x[0][e]));
    }
    if (J.nt > 1) {
        double2 d = J.d[1][0][e];
        if (J.d[1][1]) d = cadd(d, J.d[1][1][e]);
        v = csub(v, cmul(d, J.x[1][e]));
    }
    J.out[e] = v;
}

void launch_diag(hipStream_t s, const DiagJob* jobs_d, int njobs, int n) {
    k_diag<<<dim3((n + 255) / 256, njobs), 256, 0, s>>>(jobs_d, n);
}

// ---------------------------------------------------------------------------
// out = w .* in (or in ./ w), w[ix + iy Nx] = sxf[ix] syf[iy].  With the
// periodic differences A'^T = W^-1 A' W, so a transposed solve is a plain
// solve between two such scalings.
// ---------------------------------------------------------------------------
__global__ void k_wscale(double2* out, const double2* in, const double2* sxf, const double2* syf,
                         int Nx, int Ny, int inverse) {
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= (long)Nx * Ny) return;
    const int ix = (int)(k % Nx), iy = (int)(k / Nx);
    const double2 w = cmul(sxf[ix], syf[iy]);
    out[k] = inverse ? cdiv(in[k], w) : cmul(in[k], w);
}

void launch_wscale(hipStream_t s, double2* out, const double2* in, const double2* sxf,
                   const double2* syf, int Nx, int Ny, int inverse) {
    const long N = (long)Nx * Ny;
    k_wscale<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(out, in, sxf, syf, Nx, Ny, inverse);
}

}  // namespace fdfd
