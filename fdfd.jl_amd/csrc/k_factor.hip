// Block-tridiagonal factorization kernels of libfdfd (gfx950).
//
// Each grid line is a dense n x n block.  Its Schur complement S_j is formed
// from the neighbouring line's factor (k_schur) and inverted in place by
// blocked Gauss-Jordan elimination with partial pivoting.  Pivot rows are not
// swapped; they are tracked by index.  For a panel of nb columns K with
// original values V = A[:, K] and pivot rows p:
//   * k_panel    : GEPP pivot selection on the not-yet-pivoted rows of V
//                  (one workgroup, one row per thread in registers), then the
//                  inverse Vpinv of the pivot block V[p, :] = L U;
//   * k_zbuild   : Z = Vpinv * [A[p, C] | I_K]              (nb x n);
//   * k_trailing : A[i, :] = A'[i, :] - V[i, :] Z  (i not in p), A[p, :] = Z,
//                  a rank-nb complex update on MFMA (A' = A with K zeroed).
// After the last panel the stored Y satisfies S'^-1[a, b] = Y[prow[a],
// iprow[b]]; k_maps folds the row map of S' into the column map kap so that
// X_j = S_j^-1 is read as X[a, b] = Y[rho[a], kap[b]].
// The sweeps (k_sweep) apply X_j or X_j^T to a vector with that indexing.
#include "fdfd_common.h"

#include <algorithm>

namespace fdfd {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename R> __device__ inline R shfl_x(R v, int o) { return __shfl_xor(v, o, 64); }

// ---------------------------------------------------------------------------
// Schur formation: one block per row r of S'.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_schur(const SchurJob* jobs, int n) {
    using C = cplx<R>;
    const SchurJob J = jobs[blockIdx.y];
    __shared__ C row1[MAX_LINE];
    __shared__ C row2[MAX_LINE];
    const int r = blockIdx.x;
    const int a = J.Y1 ? J.rhoinv1[r] : r;  // q(r)
    if (threadIdx.x == 0) {
        J.q[r] = a;
        J.used[r] = 0;
        J.iprow[r] = -1;
    }
    if (J.Y1) {
        const C* y1 = (const C*)J.Y1 + (long)r * n;
        for (int c = threadIdx.x; c < n; c += blockDim.x) row1[c] = y1[c];
    }
    if (J.Y2) {
        const C* y2 = (const C*)J.Y2 + (long)J.rho2[a] * n;
        for (int c = threadIdx.x; c < n; c += blockDim.x) row2[c] = y2[c];
    }
    __syncthreads();
    const long si = J.si;
    const double2 dg = J.ddiag[a * si], up = J.dsup[a * si], dn = J.dsub[a * si];
    const int ap = (a + 1 == n) ? 0 : a + 1, am = (a == 0) ? n - 1 : a - 1;
    const double2 la1 = J.Y1 ? J.l1[a * si] : make_double2(0, 0);
    const double2 la2 = J.Y2 ? J.l2[a * si] : make_double2(0, 0);
    C* S = (C*)J.S + (long)r * n;
    for (int b = threadIdx.x; b < n; b += blockDim.x) {
        double2 v = make_double2(0, 0);
        if (b == a) v = cadd(v, dg);
        if (b == ap) v = cadd(v, up);
        if (b == am) v = cadd(v, dn);
        if (J.Y1) v = csub(v, cmul(cmul(la1, J.u1[b * si]), to_d(row1[J.kap1[b]])));
        if (J.Y2) v = csub(v, cmul(cmul(la2, J.u2[b * si]), to_d(row2[J.kap2[b]])));
        S[b] = from_d<C>(v);
    }
}

template <typename R>
void launch_schur(hipStream_t s, const SchurJob* jobs_d, int njobs, int n) {
    k_schur<R><<<dim3(n, njobs), 256, 0, s>>>(jobs_d, n);
}
template void launch_schur<double>(hipStream_t, const SchurJob*, int, int);
template void launch_schur<float>(hipStream_t, const SchurJob*, int, int);

// ---------------------------------------------------------------------------
// Panel: pivot selection (GEPP on unpivoted rows) + inverse of the pivot block.
// blockDim = 1024, thread t owns row t (n <= 1024).  The panel width NB is a
// compile-time constant (16, 8, 4, 2 or 1) so the row segment W[] lives in
// registers; inner loops have constant bounds and are predicated on cc.
// ---------------------------------------------------------------------------
template <typename R, int NB>
__global__ __launch_bounds__(1024) void k_panel(const LineJob* jobs, int k0) {
    using C = cplx<R>;
    const LineJob J = jobs[blockIdx.x];
    const int n = J.n;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    __shared__ R cmag[2][16];
    __shared__ int crow[2][16];
    __shared__ C cval[2][16][NB];
    __shared__ C Ls[NB][NB];
    __shared__ C Us[NB][NB];
    __shared__ C Li[NB][NB];
    __shared__ C Ui[NB][NB];
    __shared__ int prs[NB];
    __shared__ R pmags[NB];
    C* A = (C*)J.Y;
    C* Vb = (C*)J.Vbuf;
    const int r = t;
    const bool valid = r < n;
    bool act = valid && (J.used[r] == 0);
    const C zero = mk<C>(R(0), R(0));
    C W[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        W[c] = valid ? A[(long)r * n + k0 + c] : mk<C>(R(0), R(0));
        if (valid) Vb[(long)r * PMAX + c] = W[c];
    }
    int buf = 0;
#pragma unroll
    for (int cc = 0; cc < NB; ++cc) {
        R bm = act ? cabs1(W[cc]) : R(-1);
        int br = valid ? r : 0x7fffffff;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            R om = shfl_x(bm, o);
            int orr = __shfl_xor(br, o, 64);
            if (om > bm || (om == bm && orr < br)) { bm = om; br = orr; }
        }
        if (r == br) {
            cmag[buf][wv] = bm;
            crow[buf][wv] = br;
#pragma unroll
            for (int c = 0; c < NB; ++c) cval[buf][wv][c] = W[c];
        } else if (lane == 0 && br == 0x7fffffff) {
            cmag[buf][wv] = R(-1);
            crow[buf][wv] = 0x7fffffff;
        }
        __syncthreads();
        R pm = cmag[buf][0];
        int pw = 0, pr = crow[buf][0];
#pragma unroll
        for (int q = 1; q < 16; ++q) {
            R m2 = cmag[buf][q];
            int r2 = crow[buf][q];
            if (m2 > pm || (m2 == pm && r2 < pr)) { pm = m2; pw = q; pr = r2; }
        }
        if (t == 0) {
            prs[cc] = pr;
            pmags[cc] = pm;
        }
        if (t < NB) {
            C v = cval[buf][pw][t];
            Ls[cc][t] = mk<C>((t < cc) ? v.x : (t == cc ? R(1) : R(0)), (t < cc) ? v.y : R(0));
            Us[cc][t] = mk<C>((t >= cc) ? v.x : R(0), (t >= cc) ? v.y : R(0));
        }
        if (r == pr) act = false;
        if (act) {
            const C ucc = cval[buf][pw][cc];
            const C m = csel(pm > R(0), cmul(W[cc], crcp(ucc)), zero);
            W[cc] = m;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c > cc) W[c] = csub(W[c], cmul(m, cval[buf][pw][c]));
        }
        buf ^= 1;
    }
    __syncthreads();
    // Triangular inverses: column c handled by the NB lanes t = c*NB + i
    // (NB divides 64, so a column never straddles two waves).
    if (t < NB * NB) {
        const int c = t / NB, i = t % NB;
        const int gbase = (lane / NB) * NB;
        // Li = L^-1 (unit lower), forward substitution
        C x = mk<C>(i == c ? R(1) : R(0), R(0));
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            R xr = __shfl(x.x, gbase + k, 64), xi = __shfl(x.y, gbase + k, 64);
            if (i > k) x = csub(x, cmul(Ls[i][k], mk<C>(xr, xi)));
        }
        Li[i][c] = x;
        // Ui = U^-1 (upper), right-looking back substitution
        C y = mk<C>(i == c ? R(1) : R(0), R(0));
#pragma unroll
        for (int k = NB - 1; k >= 0; --k) {
            if (i == k) y = cmul(y, crcp(Us[k][k]));
            R yr = __shfl(y.x, gbase + k, 64), yi = __shfl(y.y, gbase + k, 64);
            if (i < k) y = csub(y, cmul(Us[i][k], mk<C>(yr, yi)));
        }
        Ui[i][c] = y;
    }
    __syncthreads();
    // Vpinv = U^-1 L^-1
    for (int e = t; e < NB * NB; e += blockDim.x) {
        const int i = e / NB, j = e % NB;
        C acc = zero;
        for (int k = (i > j ? i : j); k < NB; ++k) acc = cadd(acc, cmul(Ui[i][k], Li[k][j]));
        ((C*)J.Vpinv)[i * PMAX + j] = acc;
    }
    if (t < NB) {
        const int p = prs[t];
        J.piv[t] = p;
        J.prow[k0 + t] = p;
        J.iprow[p] = k0 + t;
        J.used[p] = 1;
    }
    if (t == 0) {
        double mn = J.pivstat[0], mx = J.pivstat[1];
        for (int c = 0; c < NB; ++c) {
            double v = (double)pmags[c];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
            if (!(v > 0)) J.pivstat[2] = 1.0;
        }
        J.pivstat[0] = mn;
        J.pivstat[1] = mx;
    }
}

template <typename R>
void launch_panel(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb) {
    (void)n;
    switch (nb) {
        case 16: k_panel<R, 16><<<njobs, 1024, 0, s>>>(jobs_d, k0); break;
        case 8: k_panel<R, 8><<<njobs, 1024, 0, s>>>(jobs_d, k0); break;
        case 4: k_panel<R, 4><<<njobs, 1024, 0, s>>>(jobs_d, k0); break;
        case 2: k_panel<R, 2><<<njobs, 1024, 0, s>>>(jobs_d, k0); break;
        default: k_panel<R, 1><<<njobs, 1024, 0, s>>>(jobs_d, k0); break;
    }
}
template void launch_panel<double>(hipStream_t, const LineJob*, int, int, int, int);
template void launch_panel<float>(hipStream_t, const LineJob*, int, int, int, int);

// ---------------------------------------------------------------------------
// Z = Vpinv * T, T[k][j] = delta(k, j-k0) for j in K, else A[piv[k]][j].
// One block per 64 columns.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_zbuild(const LineJob* jobs, int k0, int nb) {
    constexpr int NB = PMAX;
    using C = cplx<R>;
    const LineJob J = jobs[blockIdx.y];
    const int n = J.n;
    __shared__ C Vs[NB][NB];
    __shared__ C Ts[NB][64];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 64;
    const C* A = (const C*)J.Y;
    for (int e = t; e < NB * NB; e += 256) {
        const int i = e / NB, j = e % NB;
        Vs[i][j] = (i < nb && j < nb) ? ((const C*)J.Vpinv)[i * PMAX + j] : mk<C>(R(0), R(0));
    }
    for (int e = t; e < NB * 64; e += 256) {
        int k = e / 64, jj = e % 64, j = j0 + jj;
        C v = mk<C>(R(0), R(0));
        if (k < nb && j < n) {
            if (j >= k0 && j < k0 + nb) v = mk<C>(R(j - k0 == k ? 1 : 0), R(0));
            else v = A[(long)J.piv[k] * n + j];
        }
        Ts[k][jj] = v;
    }
    __syncthreads();
    C* Z = (C*)J.Z;
    for (int e = t; e < NB * 64; e += 256) {
        int c = e / 64, jj = e % 64, j = j0 + jj;
        if (c < nb && j < n) {
            C acc = mk<C>(R(0), R(0));
#pragma unroll
            for (int k = 0; k < NB; ++k) acc = cadd(acc, cmul(Vs[c][k], Ts[k][jj]));
            Z[(long)c * n + j] = acc;
        }
    }
}

template <typename R>
void launch_zbuild(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb) {
    dim3 grid((n + 63) / 64, njobs);
    k_zbuild<R><<<grid, 256, 0, s>>>(jobs_d, k0, nb);
}
template void launch_zbuild<double>(hipStream_t, const LineJob*, int, int, int, int);
template void launch_zbuild<float>(hipStream_t, const LineJob*, int, int, int, int);

// ---------------------------------------------------------------------------
// Trailing update on MFMA.  64x64 output tile per block, 4 waves of 32x32,
// 16x16x4 MFMA blocks, complex product as four real MFMAs:
//   re += Vr Zr - Vi Zi,  im += Vr Zi + Vi Zr.
// Operand layout of v_mfma_*_16x16x4: A[i = lane%16][k = lane/16],
// B[k = lane/16][j = lane%16]; result D[mfma_t::row(lane, v)][lane%16].
// ---------------------------------------------------------------------------
template <typename R> struct mfma_t;
// Output row of result element v held by a lane: the f32 form keeps four
// consecutive rows per lane group, the f64 form interleaves them.
template <> struct mfma_t<double> {
    using acc = d4v;
    static __device__ inline acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct mfma_t<float> {
    using acc = f4v;
    static __device__ inline acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <typename R>
__global__ __launch_bounds__(256) void k_trailing(const LineJob* jobs, int k0, int nb) {
    constexpr int NB = PMAX;
    using C = cplx<R>;
    using M = mfma_t<R>;
    using acc_t = typename M::acc;
    const LineJob J = jobs[blockIdx.z];
    const int n = J.n;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    __shared__ R Vr[64][NB + 1], Vi[64][NB + 1];
    __shared__ R Zr[NB][64 + 1], Zi[NB][64 + 1];
    const C* Vb = (const C*)J.Vbuf;
    const C* Zg = (const C*)J.Z;
    for (int e = t; e < 64 * NB; e += 256) {
        int ii = e / NB, k = e % NB, i = i0 + ii;
        C v = (i < n && k < nb) ? Vb[(long)i * PMAX + k] : mk<C>(R(0), R(0));
        Vr[ii][k] = v.x;
        Vi[ii][k] = v.y;
    }
    for (int e = t; e < 64 * NB; e += 256) {
        int k = e / 64, jj = e % 64, j = j0 + jj;
        C v = (j < n && k < nb) ? Zg[(long)k * n + j] : mk<C>(R(0), R(0));
        Zr[k][jj] = v.x;
        Zi[k][jj] = v.y;
    }
    __syncthreads();
    acc_t are[2][2], aim[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            are[bi][bj] = acc_t{0, 0, 0, 0};
            aim[bi][bj] = acc_t{0, 0, 0, 0};
        }
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < NB; kk += 4) {
        R ar[2], ai[2], br[2], bi_[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            ar[q] = Vr[wr * 32 + q * 16 + lr][kk + lk];
            ai[q] = Vi[wr * 32 + q * 16 + lr][kk + lk];
            br[q] = Zr[kk + lk][wc * 32 + q * 16 + lr];
            bi_[q] = Zi[kk + lk][wc * 32 + q * 16 + lr];
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                are[bi][bj] = M::run(ar[bi], br[bj], are[bi][bj]);
                are[bi][bj] = M::run(-ai[bi], bi_[bj], are[bi][bj]);
                aim[bi][bj] = M::run(ar[bi], bi_[bj], aim[bi][bj]);
                aim[bi][bj] = M::run(ai[bi], br[bj], aim[bi][bj]);
            }
    }
    C* A = (C*)J.Y;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ii = wr * 32 + bi * 16 + M::row(lane, v);
            const int i = i0 + ii;
            if (i >= n) continue;
            const int slot = J.iprow[i] - k0;
            const bool pivrow = (slot >= 0 && slot < nb);
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int jj = wc * 32 + bj * 16 + lr;
                const int j = j0 + jj;
                if (j >= n) continue;
                C out;
                if (pivrow) {
                    out = Zg[(long)slot * n + j];
                } else {
                    C old = (j >= k0 && j < k0 + nb) ? mk<C>(R(0), R(0)) : A[(long)i * n + j];
                    out = mk<C>(old.x - are[bi][bj][v], old.y - aim[bi][bj][v]);
                }
                A[(long)i * n + j] = out;
            }
        }
}

template <typename R>
void launch_trailing(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb) {
    dim3 grid((n + 63) / 64, (n + 63) / 64, njobs);
    k_trailing<R><<<grid, 256, 0, s>>>(jobs_d, k0, nb);
}
template void launch_trailing<double>(hipStream_t, const LineJob*, int, int, int, int);
template void launch_trailing<float>(hipStream_t, const LineJob*, int, int, int, int);

// ---------------------------------------------------------------------------
// kapinv[c] = q[prow[c]], kap[kapinv[c]] = c
// ---------------------------------------------------------------------------
__global__ void k_maps(const LineJob* jobs) {
    const LineJob J = jobs[blockIdx.y];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= J.n) return;
    int b = J.q[J.prow[c]];
    J.kapinv[c] = b;
    J.kap[b] = c;
}
void launch_maps(hipStream_t s, const LineJob* jobs_d, int njobs, int n) {
    k_maps<<<dim3((n + 255) / 256, njobs), 256, 0, s>>>(jobs_d);
}

// ---------------------------------------------------------------------------
// Sweep step.  Non-transposed: w[c] = v[kapinv[c]], z = Y w,
// out[rhoinv[r]] = z[r] (forward/middle) or g - z[r] (backward).
// Transposed: w[r] = v[rhoinv[r]], z = Y^T w, out[kapinv[c]] = z[c] (or g - z).
//   forward : v = b - c1 .* nb1           (nb1 may be null)
//   middle  : v = b - c1 .* nb1 - c2 .* nb2
//   backward: v = c1 .* nb1
// ---------------------------------------------------------------------------
template <typename R>
__device__ inline void build_w(const SweepJob& J, const int* map, cplx<R>* w, int n) {
    using C = cplx<R>;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const long ea = (long)map[c] * J.si;
        double2 v;
        if (J.mode == 1) {
            v = cmul(J.c1[ea], J.nb1[ea]);
        } else {
            v = J.b[ea];
            if (J.nb1) v = csub(v, cmul(J.c1[ea], J.nb1[ea]));
            if (J.mode == 2 && J.nb2) v = csub(v, cmul(J.c2[ea], J.nb2[ea]));
        }
        w[c] = from_d<C>(v);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_sweep(const SweepJob* jobs, int n) {
    using C = cplx<R>;
    const SweepJob J = jobs[blockIdx.y];
    __shared__ C w[MAX_LINE];
    build_w<R>(J, J.kapinv, w, n);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rows_per_block = (n + gridDim.x - 1) / gridDim.x;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    const C* Y = (const C*)J.Y;
    for (int r = r0 + wv; r < r1; r += 4) {
        const C* yr = Y + (long)r * n;
        R sr = 0, sim = 0;
        for (int c = lane; c < n; c += 64) {
            C y = yr[c], x = w[c];
            sr += y.x * x.x - y.y * x.y;
            sim += y.x * x.y + y.y * x.x;
        }
        for (int o = 32; o > 0; o >>= 1) {
            sr += shfl_x(sr, o);
            sim += shfl_x(sim, o);
        }
        if (lane == 0) {
            const long ea = (long)J.rhoinv[r] * J.si;
            double2 z = make_double2((double)sr, (double)sim);
            J.out[ea] = (J.mode == 1) ? csub(J.g[ea], z) : z;
        }
    }
}

// transposed: blocks (column group of 64) x (row split) write partial sums;
// k_sweep_t_finish adds the TSPLIT partials of each column in a fixed order.
constexpr int TSPLIT = 16;
template <typename R>
__global__ __launch_bounds__(256) void k_sweep_t(const SweepJob* jobs, int n, double2* partial) {
    using C = cplx<R>;
    const int jid = blockIdx.z;
    const SweepJob J = jobs[jid];
    __shared__ C w[MAX_LINE];
    __shared__ double2 red[4][64];
    build_w<R>(J, J.rhoinv, w, n);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    const int rows = (n + TSPLIT - 1) / TSPLIT;
    const int r0 = blockIdx.y * rows, r1 = min(n, r0 + rows);
    const C* Y = (const C*)J.Y;
    R sr = 0, sim = 0;
    if (c < n) {
        for (int r = r0 + wv; r < r1; r += 4) {
            C y = Y[(long)r * n + c], x = w[r];
            sr += y.x * x.x - y.y * x.y;
            sim += y.x * x.y + y.y * x.x;
        }
    }
    red[wv][lane] = make_double2((double)sr, (double)sim);
    __syncthreads();
    if (wv == 0 && c < n) {
        double2 s = red[0][lane];
        for (int q = 1; q < 4; ++q) s = cadd(s, red[q][lane]);
        partial[((long)jid * TSPLIT + blockIdx.y) * MAX_LINE + c] = s;
    }
}

__global__ void k_sweep_t_finish(const SweepJob* jobs, int n, const double2* partial) {
    const int jid = blockIdx.y;
    const SweepJob J = jobs[jid];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double2 s = make_double2(0, 0);
    for (int q = 0; q < TSPLIT; ++q) s = cadd(s, partial[((long)jid * TSPLIT + q) * MAX_LINE + c]);
    const long ea = (long)J.kapinv[c] * J.si;
    J.out[ea] = (J.mode == 1) ? csub(J.g[ea], s) : s;
}

template <typename R>
void launch_sweep(hipStream_t s, const SweepJob* jobs_d, int njobs, int n, int transpose,
                  double2* partial) {
    if (!transpose) {
        int nblk = std::max(1, std::min(256 / std::max(1, njobs), (n + 3) / 4));
        k_sweep<R><<<dim3(nblk, njobs), 256, 0, s>>>(jobs_d, n);
    } else {
        k_sweep_t<R><<<dim3((n + 63) / 64, TSPLIT, njobs), 256, 0, s>>>(jobs_d, n, partial);
        k_sweep_t_finish<<<dim3((n + 255) / 256, njobs), 256, 0, s>>>(jobs_d, n, partial);
    }
}
template void launch_sweep<double>(hipStream_t, const SweepJob*, int, int, int, double2*);
template void launch_sweep<float>(hipStream_t, const SweepJob*, int, int, int, double2*);

}  // namespace fdfd
