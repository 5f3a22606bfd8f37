// Several right-hand sides per launch: the reduced-solve steps and the FP64
// Krylov building blocks of a lockstep multi-RHS FGMRES.
//
// The reference solves one source per solve() call and refactors per
// frequency (src/solver/driven.jl:11-55).  Here the sources that share a
// factor are solved together: every step of the reduced solve streams a
// factor block once for a chunk of KC right-hand sides (instead of once per
// source), and each Arnoldi step does its dot products, updates and norms
// for all sources in one launch and one host round trip.
//
// Vector layout: right-hand side r of a buffer starts at r * rs (rs >= N);
// only the sources listed in `ridx` (na of them) are touched, so converged
// sources drop out of the launches.
#include <hip/hip_runtime.h>

#include "fdfd_common.h"
#include "k_device.h"
#include "tri_qr.h"

namespace fdfd {

// ---------------------------------------------------------------------------
// reduced solve, dense matrix-vector step for KC right-hand sides at once:
// out_r[omap(i)] = [base_r] - sum_t M_t[i, :] x_{t,r}[imap_t(:)]
// Each workgroup owns a row range of one job and one chunk of KC sources;
// the sources' gathered inputs sit in LDS, and every matrix entry loaded is
// used for all KC of them.
// ---------------------------------------------------------------------------
template <typename R, int KC>
__global__ __launch_bounds__(256) void k_mv_multi(const MvJob* jobs, int n, long rs,
                                                  const int* ridx, int na) {
    using C = cplx<R>;
    extern __shared__ __align__(16) unsigned char mv_lds[];
    C* w = (C*)mv_lds;  // [term][KC][n]
    const MvJob J = jobs[blockIdx.y];
    const int q0 = blockIdx.z * KC;
    const int kc = min(KC, na - q0);
    const long si = J.si;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        for (int t = 0; t < J.nt; ++t) {
            const int s = J.imap[t] ? J.imap[t][c] : c;
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                if (q < kc) {
                    const long r = ridx[q0 + q];
                    w[((long)t * KC + q) * n + c] = from_d<C>(J.x[t][(long)s * si + r * rs]);
                }
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rows_per_block = (n + gridDim.x - 1) / gridDim.x;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    for (int row = r0 + 2 * wv; row < r1; row += 8) {
        const bool two = (row + 1) < r1;
        const int rb = two ? row + 1 : row;
        R ar[KC], ai[KC], br[KC], bi[KC];
#pragma unroll
        for (int q = 0; q < KC; ++q) ar[q] = ai[q] = br[q] = bi[q] = R(0);
        for (int t = 0; t < J.nt; ++t) {
            const C* Ma = (const C*)J.M[t] + (long)row * n;
            const C* Mb = (const C*)J.M[t] + (long)rb * n;
            const C* wt = w + (long)t * KC * n;
            for (int c = lane; c < n; c += 64) {
                const C a = gload(Ma + c), b = gload(Mb + c);
#pragma unroll
                for (int q = 0; q < KC; ++q) {
                    const C x = wt[(long)q * n + c];
                    ar[q] += a.x * x.x - a.y * x.y;
                    ai[q] += a.x * x.y + a.y * x.x;
                    br[q] += b.x * x.x - b.y * x.y;
                    bi[q] += b.x * x.y + b.y * x.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < KC; ++q) {
            for (int o = 32; o > 0; o >>= 1) {
                ar[q] += shfl_x(ar[q], o);
                ai[q] += shfl_x(ai[q], o);
                br[q] += shfl_x(br[q], o);
                bi[q] += shfl_x(bi[q], o);
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < KC; ++q) {
                if (q < kc) {
                    const long ro = (long)ridx[q0 + q] * rs;
                    long e = (long)(J.omap ? J.omap[row] : row) * si + ro;
                    double2 z = make_double2((double)ar[q], (double)ai[q]);
                    J.out[e] = J.base ? csub(J.base[e], z) : z;
                    if (two) {
                        e = (long)(J.omap ? J.omap[row + 1] : row + 1) * si + ro;
                        z = make_double2((double)br[q], (double)bi[q]);
                        J.out[e] = J.base ? csub(J.base[e], z) : z;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The same step on the matrix cores, for a chunk of up to MVQ = 16
// right-hand sides: Y (n x 16) = M (n x n) X (n x 16) is a skinny complex
// GEMM, run as v_mfma 16x16x4 products (f32 for complex64 factors, f64 for
// complex128).  Two launches:
//  - k_mv_pack gathers each job's inputs once (through imap, converted to the
//    factor's precision) into a packed panel Xp[job][term][k][16], zero past
//    n and past the chunk's sources, so that
//  - k_mv_mfma streams the factor blocks with no gather, barrier or LDS in its
//    loop: a workgroup of 4 waves owns 16 RW rows of one job; wave w takes row
//    tile w % RW and K part w / RW of KS = 4 / RW parts (chunks of MVK
//    columns, chunk c in part c % KS), issuing the next chunk's A and B
//    loads before the current chunk's products.  The lane's A operand for
//    K-step j is M[row = lane % 16][k0 + lane / 16 + 4 j] (four lanes of a
//    row read 4 adjacent entries per load), its B operand the matching
//    packed row (a wave reads 512 contiguous bytes per K-step).  Re and Im go
//    to separate accumulators for each of the four real products; the K
//    parts are summed through LDS at the end.
// Every factor entry is loaded once per chunk of 16 sources, and the packed
// panel (16 sources x 8 or 16 B per column) is read from cache by each row
// tile.  RW = 4 is used for steps whose jobs fill the chip, RW = 2 or 1 split
// the K range for short steps, whose time is latency.
// ---------------------------------------------------------------------------
constexpr int MVQ = 16;   // right-hand sides per chunk of sources (the MFMA's N)
constexpr int MVP = 64;   // packed panels are padded to a multiple of this many columns

template <typename R>
__global__ __launch_bounds__(256) void k_mv_pack(const MvJob* jobs, int n, int npad, long rs,
                                                 const int* ridx, int na, cplx<R>* Xp) {
    using C = cplx<R>;
    __shared__ C tile[MVP][MVQ + 1];
    const MvJob J = jobs[blockIdx.y];
    const int t = blockIdx.z;
    if (t >= J.nt) return;
    const int k0 = blockIdx.x * MVP;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kc = min(MVQ, na);
    const double2* x = t ? J.x[1] : J.x[0];
    const int* im = t ? J.imap[1] : J.imap[0];
    const int k = k0 + lane;
    const long sk = k < n ? (long)(im ? gload_s(im + k) : k) * J.si : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int q = 4 * i + wv;
        tile[lane][q] = (k < n && q < kc) ? from_d<C>(gload(x + sk + (long)gload_s(ridx + q) * rs))
                                          : mk<C>(0, 0);
    }
    __syncthreads();
    C* dst = Xp + ((long)(2 * blockIdx.y + t) * npad + k0) * MVQ;
#pragma unroll
    for (int i = 0; i < MVP * MVQ / 256; ++i) {
        const int e = i * 256 + threadIdx.x;
        gstore(dst + e, tile[e / MVQ][e % MVQ]);
    }
}

template <typename R, int RW>
__global__ __launch_bounds__(256) void k_mv_mfma(const MvJob* jobs, int n, int npad, long rs,
                                                 const int* ridx, int na, const cplx<R>* Xp) {
    using C = cplx<R>;
    using MF = mfma_t<R>;
    using acc_t = typename MF::acc;
    constexpr int KS = 4 / RW;
    constexpr int MVK = sizeof(R) == 4 ? 64 : 32;  // columns per chunk
    constexpr int KJ = MVK / 4;                    // K-steps per chunk
    __shared__ R red[KS > 1 ? (KS - 1) * RW * 64 * 8 : 1];
    const int job = blockIdx.y;
    const MvJob J = jobs[job];
    const int kc = min(MVQ, na);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rt = wv % RW, kp = wv / RW;
    const int kq = lane >> 4;
    const int row_base = blockIdx.x * 16 * RW + rt * 16;
    const int rowA = row_base + (lane & 15);
    const bool rowok = rowA < n;
    const int nkb = npad / MVK;
    const int nch = J.nt * nkb;
    const C* xb = Xp + (long)2 * job * npad * MVQ + (lane & 15);
    auto load = [&](int ch, C* A, C* B) {
        const int t = ch / nkb, k0 = (ch - t * nkb) * MVK + kq;
        const C* row = (const C*)(t ? J.M[1] : J.M[0]) + (long)(rowok ? rowA : 0) * n;
        const C* xc = xb + ((long)t * npad + k0) * MVQ;
#pragma unroll
        for (int j = 0; j < KJ; ++j) {
            const int k = k0 + 4 * j;
            A[j] = (rowok && k < n) ? gload(row + k) : mk<C>(0, 0);
            B[j] = gload(xc + 4 * j * MVQ);
        }
    };
    C a[KJ], b[KJ], an[KJ], bn[KJ];
    acc_t crr = {0, 0, 0, 0}, cii = {0, 0, 0, 0}, cri = {0, 0, 0, 0}, cir = {0, 0, 0, 0};
    if (kp < nch) load(kp, a, b);
    for (int ch = kp; ch < nch; ch += KS) {
        const bool more = ch + KS < nch;
        if (more) load(ch + KS, an, bn);
#pragma unroll
        for (int j = 0; j < KJ; ++j) {
            crr = MF::run(a[j].x, b[j].x, crr);
            cii = MF::run(-a[j].y, b[j].y, cii);
            cri = MF::run(a[j].x, b[j].y, cri);
            cir = MF::run(a[j].y, b[j].x, cir);
        }
        if (more) {
#pragma unroll
            for (int j = 0; j < KJ; ++j) {
                a[j] = an[j];
                b[j] = bn[j];
            }
        }
    }
    acc_t yr = crr + cii, yi = cri + cir;
    if constexpr (KS > 1) {
        // parts 1 .. KS-1 leave their sums in LDS, part 0 adds them
        if (kp > 0) {
            R* d = red + ((long)((kp - 1) * RW + rt) * 64 + lane) * 8;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                d[v] = yr[v];
                d[4 + v] = yi[v];
            }
        }
        __syncthreads();
        if (kp > 0) return;
#pragma unroll
        for (int p = 1; p < KS; ++p) {
            const R* d = red + ((long)((p - 1) * RW + rt) * 64 + lane) * 8;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                yr[v] += d[v];
                yi[v] += d[4 + v];
            }
        }
    }
    const int q = lane & 15;
    if (q >= kc) return;
    const long ro = (long)gload_s(ridx + q) * rs;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = row_base + MF::row(lane, v);
        if (row >= n) continue;
        const long e = (long)(J.omap ? gload_s(J.omap + row) : row) * J.si + ro;
        const double2 z = make_double2((double)yr[v], (double)yi[v]);
        J.out[e] = J.base ? csub(gload(J.base + e), z) : z;
    }
}

size_t mv_pack_bytes(int njobs, int n, int fp64) {
    const size_t npad = (size_t)(n + MVP - 1) / MVP * MVP;
    return (size_t)njobs * 2 * npad * MVQ * (fp64 ? sizeof(double2) : sizeof(float2));
}

template <typename R>
static int mv_kc(int n) {
    // LDS for two terms of KC gathered vectors: at most 128 KB
    const int per = 2 * n * (int)sizeof(cplx<R>);
    int kc = 8;
    while (kc > 1 && kc * per > 128 * 1024) kc >>= 1;
    return kc;
}

template <typename R>
int mv_multi_prepare(int n) {
    const int kc = mv_kc<R>(n);
    const int lds = 2 * kc * n * (int)sizeof(cplx<R>);
    hipError_t e = hipSuccess;
    switch (kc) {
        case 8: e = hipFuncSetAttribute((const void*)k_mv_multi<R, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, lds); break;
        case 4: e = hipFuncSetAttribute((const void*)k_mv_multi<R, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, lds); break;
        case 2: e = hipFuncSetAttribute((const void*)k_mv_multi<R, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, lds); break;
        default: e = hipFuncSetAttribute((const void*)k_mv_multi<R, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds); break;
    }
    return (int)e;
}
template int mv_multi_prepare<double>(int);
template int mv_multi_prepare<float>(int);

template <typename R>
void launch_mv_multi(hipStream_t s, const MvJob* jobs_d, int njobs, int n, long rs, const int* ridx,
                     int na, bool mfma, void* pack) {
    if (mfma) {
        // sources in chunks of MVQ; 64-row workgroups when they give at
        // least two per CU, else the K range split two or four ways
        const int npad = (n + MVP - 1) / MVP * MVP;
        const long w64 = (long)njobs * ((n + 63) / 64), w32 = (long)njobs * ((n + 31) / 32);
        const int rw = w64 >= 512 ? 4 : w32 >= 512 ? 2 : 1;
        cplx<R>* Xp = (cplx<R>*)pack;
        for (int q0 = 0; q0 < na; q0 += MVQ) {
            const int nc = std::min(MVQ, na - q0);
            k_mv_pack<R><<<dim3(npad / MVP, njobs, 2), 256, 0, s>>>(jobs_d, n, npad, rs, ridx + q0,
                                                                   nc, Xp);
            const dim3 grid((n + 16 * rw - 1) / (16 * rw), njobs);
            if (rw == 4)
                k_mv_mfma<R, 4><<<grid, 256, 0, s>>>(jobs_d, n, npad, rs, ridx + q0, nc, Xp);
            else if (rw == 2)
                k_mv_mfma<R, 2><<<grid, 256, 0, s>>>(jobs_d, n, npad, rs, ridx + q0, nc, Xp);
            else
                k_mv_mfma<R, 1><<<grid, 256, 0, s>>>(jobs_d, n, npad, rs, ridx + q0, nc, Xp);
        }
        return;
    }
    const int kc = mv_kc<R>(n);
    const int nch = (na + kc - 1) / kc;
    int nblk = (4096 + njobs * nch - 1) / (njobs * nch);
    nblk = std::max(1, std::min(nblk, (n + 7) / 8));
    const dim3 grid(nblk, njobs, nch);
    const size_t lds = 2 * (size_t)kc * n * sizeof(cplx<R>);
    switch (kc) {
        case 8: k_mv_multi<R, 8><<<grid, 256, lds, s>>>(jobs_d, n, rs, ridx, na); break;
        case 4: k_mv_multi<R, 4><<<grid, 256, lds, s>>>(jobs_d, n, rs, ridx, na); break;
        case 2: k_mv_multi<R, 2><<<grid, 256, lds, s>>>(jobs_d, n, rs, ridx, na); break;
        default: k_mv_multi<R, 1><<<grid, 256, lds, s>>>(jobs_d, n, rs, ridx, na); break;
    }
}
template void launch_mv_multi<double>(hipStream_t, const MvJob*, int, int, long, const int*, int,
                                      bool, void*);
template void launch_mv_multi<float>(hipStream_t, const MvJob*, int, int, long, const int*, int,
                                     bool, void*);

// first-level (diagonal) step, one source per grid row
__global__ void k_diag_multi(const DiagJob* jobs, int n, long rs, const int* ridx) {
    const DiagJob J = jobs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long e = (long)i * J.si;
    const long ro = (long)ridx[blockIdx.z] * rs;
    double2 v = J.x0[e + ro];
    if (J.nt > 0) {
        double2 d = J.d[0][0][e];
        if (J.d[0][1]) d = cadd(d, J.d[0][1][e]);
        v = csub(v, cmul(d, J.x[0][e + ro]));
    }
    if (J.nt > 1) {
        double2 d = J.d[1][0][e];
        if (J.d[1][1]) d = cadd(d, J.d[1][1][e]);
        v = csub(v, cmul(d, J.x[1][e + ro]));
    }
    J.out[e + ro] = v;
}

void launch_diag_multi(hipStream_t s, const DiagJob* jobs_d, int njobs, int n, long rs,
                       const int* ridx, int na) {
    k_diag_multi<<<dim3((n + 255) / 256, njobs, na), 256, 0, s>>>(jobs_d, n, rs, ridx);
}

// ---------------------------------------------------------------------------
// First-level line solves for several sources: the recurrences of k_trisolve
// (y = Q in, then out = R^-1 y, tri_qr.h) run by lanes 0 .. kt-1 of wave 0,
// one source each, in the same arithmetic order; wave 1 stages the
// rotations and rows of R once for all of them (broadcast LDS reads).  Each
// source's vector sits in LDS with a padded stride (no bank conflicts).
// ---------------------------------------------------------------------------
constexpr int TM_CH = 128;   // steps per staged chunk
constexpr int TM_STEP = 10;  // doubles per step (6 rotation, 10 row of R)
constexpr int TM_U = 8;      // steps whose LDS operands are loaded together

static int tm_kt(int n) {
    // vectors (kt x (n + 1) complex) + staging (2 x TM_CH x TM_STEP doubles) <= 150 KB
    const size_t stage = 2 * TM_CH * TM_STEP * sizeof(double);
    int kt = 16;
    while (kt > 1 && kt * (size_t)(n + 1) * sizeof(double2) + stage > 150 * 1024) --kt;
    return kt;
}
static size_t tm_lds(int n, int kt) {
    return kt * (size_t)(n + 1) * sizeof(double2) + 2 * TM_CH * TM_STEP * sizeof(double);
}

__global__ __launch_bounds__(128) void k_trisolve_multi(const TrsJob* jobs, int n, long rs,
                                                        const int* ridx, int na, int kt) {
    extern __shared__ __align__(16) double tmm[];
    const int ldl = n + 1;
    double2* vall = (double2*)tmm;
    double* buf = tmm + 2 * (size_t)kt * ldl;
    const TrsJob J = jobs[blockIdx.x];
    const int q0 = blockIdx.y * kt;
    const int nk = min(kt, na - q0);
    const int t = threadIdx.x;
    const bool comp = t < nk, loader = t >= 64;
    const int ll = t - 64;
    auto stage = [&](int b, long src0, int cnt) {
        constexpr int PER = TM_CH * TM_STEP / 64;
        double r[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 64 + ll;
            r[q] = gload_s(J.w + src0 + (e < cnt ? e : 0));
        }
        double* d = buf + b * TM_CH * TM_STEP;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 64 + ll;
            if (e < cnt) d[e] = r[q];
        }
    };
    // gather the nk input vectors
    for (int q = 0; q < nk; ++q) {
        const long ro = (long)ridx[q0 + q] * rs;
        for (int k = t; k < n; k += 128) vall[(long)q * ldl + k] = gload(J.in + (long)k * J.si + ro);
    }
    double2* v = vall + (long)(comp ? t : 0) * ldl;
    using tri::cz;
    auto ld = [&](int k) { return tri::mkz(v[k].x, v[k].y); };
    auto st = [&](int k, cz z) { v[k] = make_double2(z.re, z.im); };
    const int n1 = n - 1;
    const int nc1 = (n1 + TM_CH - 1) / TM_CH;
    if (loader && n1 > 0) stage(0, 0, tri::ROT * min(TM_CH, n1));
    __syncthreads();
    cz carry = tri::mkz(0, 0), tail = tri::mkz(0, 0);
    if (comp && n > 1) {
        carry = ld(0);
        tail = ld(n - 1);
    }
    for (int c = 0; c < nc1; ++c) {
        const int k0 = c * TM_CH, len = min(TM_CH, n1 - k0);
        if (loader && c + 1 < nc1) {
            const int k1 = k0 + TM_CH;
            stage((c + 1) & 1, (long)tri::ROT * k1, tri::ROT * min(TM_CH, n1 - k1));
        }
        if (comp) {
            const double* rb = buf + (c & 1) * TM_CH * TM_STEP;
            // general steps k < n - 2; the final rotation (k = n - 2) is
            // peeled off, so batches of TM_U steps run as straight-line code
            // with their operands loaded first
            const int lenG = max(0, min(len, n - 2 - k0));
            auto gstep = [&](const double* r, cz nxt, int k) {
                tri::rot(r[0], tri::mkz(r[1], r[2]), carry, nxt);
                tri::rot(r[3], tri::mkz(r[4], r[5]), carry, tail);
                st(k, carry);
                carry = nxt;
            };
            int i0 = 0;
            for (; i0 + TM_U <= lenG; i0 += TM_U) {
                double rr[TM_U][tri::ROT];
                cz nx[TM_U];
#pragma unroll
                for (int u = 0; u < TM_U; ++u) {
#pragma unroll
                    for (int j = 0; j < tri::ROT; ++j) rr[u][j] = rb[tri::ROT * (i0 + u) + j];
                    nx[u] = ld(k0 + i0 + u + 1);
                }
#pragma unroll
                for (int u = 0; u < TM_U; ++u) gstep(rr[u], nx[u], k0 + i0 + u);
            }
            for (; i0 < lenG; ++i0) gstep(rb + tri::ROT * i0, ld(k0 + i0 + 1), k0 + i0);
            if (k0 + len - 1 == n - 2) {
                const double* r = rb + tri::ROT * (len - 1);
                tri::rot(r[0], tri::mkz(r[1], r[2]), carry, tail);
                st(n - 2, carry);
                st(n - 1, tail);
            }
        }
        __syncthreads();
    }
    const long rbase = (long)tri::ROT * n;
    const int nc2 = (n + TM_CH - 1) / TM_CH;
    if (loader) {
        const int lo = max(0, n - TM_CH);
        stage(0, rbase + (long)tri::RROW * lo, tri::RROW * (n - lo));
    }
    __syncthreads();
    cz x1 = tri::mkz(0, 0), x2 = tri::mkz(0, 0), xA = tri::mkz(0, 0), xB = tri::mkz(0, 0);
    for (int c = 0; c < nc2; ++c) {
        const int hi = n - c * TM_CH, lo = max(0, hi - TM_CH);
        if (loader && c + 1 < nc2) {
            const int lo2 = max(0, lo - TM_CH);
            stage((c + 1) & 1, rbase + (long)tri::RROW * lo2, tri::RROW * (lo - lo2));
        }
        if (comp) {
            const double* rb = buf + (c & 1) * TM_CH * TM_STEP;
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
            for (; kt - (TM_U - 1) >= lo; kt -= TM_U) {
                double rr[TM_U][tri::RROW];
                cz yy[TM_U];
#pragma unroll
                for (int u = 0; u < TM_U; ++u) {
#pragma unroll
                    for (int j = 0; j < tri::RROW; ++j) rr[u][j] = row(kt - u)[j];
                    yy[u] = ld(kt - u);
                }
#pragma unroll
                for (int u = 0; u < TM_U; ++u) gstep(rr[u], yy[u], kt - u);
            }
            for (; kt >= lo; --kt) gstep(row(kt), ld(kt), kt);
        }
        __syncthreads();
    }
    for (int q = 0; q < nk; ++q) {
        const long ro = (long)ridx[q0 + q] * rs;
        for (int k = t; k < n; k += 128) gstore(J.out + (long)k * J.si + ro, vall[(long)q * ldl + k]);
    }
}

int trisolve_multi_prepare(int n) {
    return (int)hipFuncSetAttribute((const void*)k_trisolve_multi,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)tm_lds(n, tm_kt(n)));
}

void launch_trisolve_multi(hipStream_t s, const TrsJob* jobs_d, int njobs, int n, long rs,
                           const int* ridx, int na) {
    const int kt = tm_kt(n);
    const dim3 grid(njobs, (na + kt - 1) / kt);
    k_trisolve_multi<<<grid, 128, tm_lds(n, kt), s>>>(jobs_d, n, rs, ridx, na, kt);
}

// ---------------------------------------------------------------------------
// Krylov building blocks over the active sources (grid row y -> source
// ridx[y]).  Basis vector j of source r sits at V + j * ldv + r * rs.
// ---------------------------------------------------------------------------
constexpr int MB_MAXK = 32;

// partial[y][blk][i] = sum over this block's cells of conj(V_{i,r}) w_r
__global__ __launch_bounds__(256) void k_multidot_b(const double2* V, long ldv, long rs, int k,
                                                    const double2* w, long wrs, long n,
                                                    double2* partial, const int* ridx) {
    __shared__ double2 red[4][MB_MAXK];
    const long r = ridx[blockIdx.y];
    const double2* Vr = V + r * rs;
    const double2* wr = w + r * wrs;
    double2 acc[MB_MAXK];
#pragma unroll
    for (int i = 0; i < MB_MAXK; ++i) acc[i] = make_double2(0, 0);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const double2 we = wr[e];
#pragma unroll
        for (int i = 0; i < MB_MAXK; ++i)
            if (i < k) acc[i] = cadd(acc[i], cconjmul(Vr[i * ldv + e], we));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < MB_MAXK; ++i) {
        if (i < k) {
            double re = acc[i].x, im = acc[i].y;
            for (int o = 32; o > 0; o >>= 1) {
                re += __shfl_xor(re, o, 64);
                im += __shfl_xor(im, o, 64);
            }
            if (lane == 0) red[wv][i] = make_double2(re, im);
        }
    }
    __syncthreads();
    if (threadIdx.x < k) {
        double2 s = red[0][threadIdx.x];
        for (int q = 1; q < 4; ++q) s = cadd(s, red[q][threadIdx.x]);
        partial[((long)blockIdx.y * gridDim.x + blockIdx.x) * k + threadIdx.x] = s;
    }
}

// out[y][i] = sum_blk partial[y][blk][i]: thread t adds column t % 32 over
// rows t / 32 + 32 q (fixed order), then the 32 row groups in LDS
__global__ __launch_bounds__(1024) void k_reduce_partials_b(const double2* partial, int nblk, int k,
                                                            double2* out) {
    __shared__ double2 acc[32][32];
    const int y = blockIdx.x;
    const double2* P = partial + (long)y * nblk * k;
    const int c = threadIdx.x % 32, grp = threadIdx.x / 32;
    for (int c0 = 0; c0 < k; c0 += 32) {
        const int i = c0 + c;
        double2 s = make_double2(0, 0);
        if (i < k)
            for (int b = grp; b < nblk; b += 32) s = cadd(s, P[b * (long)k + i]);
        acc[grp][c] = s;
        __syncthreads();
        if (grp == 0 && i < k) {
            double2 t = acc[0][c];
            for (int g2 = 1; g2 < 32; ++g2) t = cadd(t, acc[g2][c]);
            out[(long)y * k + i] = t;
        }
        __syncthreads();
    }
}

void launch_multidot_b(hipStream_t s, const double2* V, long ldv, long rs, int k, const double2* w,
                       long wrs, long n, double2* partial, double2* out, int nblk,
                       const int* ridx, int na) {
    k_multidot_b<<<dim3(nblk, na), 256, 0, s>>>(V, ldv, rs, k, w, wrs, n, partial, ridx);
    k_reduce_partials_b<<<na, 1024, 0, s>>>(partial, nblk, k, out);
}

// w_r += sign * sum_{i < kcnt[y]} h[y][i] V_{i,r}   (h has kstride entries per source)
__global__ void k_gemv_update_b(double2* w, long wrs, const double2* V, long ldv, long rs,
                                const int* kcnt, const double2* h, int kstride, double sign,
                                long n, const int* ridx) {
    __shared__ double2 hs[64];
    const int y = blockIdx.y;
    const long r = ridx[y];
    const int k = kcnt[y];
    if (threadIdx.x < k) hs[threadIdx.x] = cscale(h[(long)y * kstride + threadIdx.x], sign);
    __syncthreads();
    double2* wr = w + r * wrs;
    const double2* Vr = V + r * rs;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        double2 acc = wr[e];
        for (int i = 0; i < k; ++i) acc = cadd(acc, cmul(hs[i], Vr[i * ldv + e]));
        wr[e] = acc;
    }
}

void launch_gemv_update_b(hipStream_t s, double2* w, long wrs, const double2* V, long ldv, long rs,
                          const int* kcnt, const double2* h, int kstride, double sign, long n,
                          const int* ridx, int na) {
    const unsigned nb = (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 2048 / na + 1));
    k_gemv_update_b<<<dim3(nb, na), 256, 0, s>>>(w, wrs, V, ldv, rs, kcnt, h, kstride, sign, n, ridx);
}

// out[y] = ||x_r||^2
__global__ __launch_bounds__(256) void k_norm2_b(const double2* x, long rs, long n, double* partial,
                                                 const int* ridx) {
    __shared__ double red[4];
    const double2* xr = x + (long)ridx[blockIdx.y] * rs;
    double acc = 0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x)
        acc += cnorm2(xr[e]);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ __launch_bounds__(256) void k_reduce_real_b(const double* partial, int nblk, double* out) {
    __shared__ double acc[256];
    const double* P = partial + (long)blockIdx.x * nblk;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += P[b];
    acc[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) acc[threadIdx.x] += acc[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = acc[0];
}
void launch_norm2_b(hipStream_t s, const double2* x, long rs, long n, double* partial, double* out,
                    int nblk, const int* ridx, int na) {
    k_norm2_b<<<dim3(nblk, na), 256, 0, s>>>(x, rs, n, partial, ridx);
    k_reduce_real_b<<<na, 256, 0, s>>>(partial, nblk, out);
}

// y_r = a[y] * x_r
__global__ void k_scale_b(double2* y, long yrs, const double2* x, long xrs, const double2* a, long n,
                          const int* ridx) {
    const long r = ridx[blockIdx.y];
    const double2 s = a[blockIdx.y];
    double2* yr = y + r * yrs;
    const double2* xr = x + r * xrs;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x)
        yr[e] = cmul(s, xr[e]);
}
void launch_scale_b(hipStream_t s, double2* y, long yrs, const double2* x, long xrs, const double2* a,
                    long n, const int* ridx, int na) {
    const unsigned nb = (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 2048 / na + 1));
    k_scale_b<<<dim3(nb, na), 256, 0, s>>>(y, yrs, x, xrs, a, n, ridx);
}

}  // namespace fdfd
