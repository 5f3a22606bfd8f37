// Operator-side kernels of libfdfd (all FP64):
//   * SC-PML inverse stretch factors        (src/pml.jl:1-63)
//   * 5-point coefficient assembly TM / TE   (src/solver/driven.jl:21-46,
//                                             src/grid.jl:128-162)
//   * matrix-free apply / residual, A and A^T
//   * BLAS-1 reductions for the FP64 Krylov refinement
//   * field recovery (driven.jl:40-41,50-51; eigen.jl:90-91,108-109)
//   * adjoint gradient product
// Field-shaped arrays are Nx x Ny complex, x fastest.  These kernels are
// HBM-bound streaming kernels: one thread per cell, coalesced along x.
#include "fdfd_common.h"

#include <algorithm>

namespace fdfd {

// ---------------------------------------------------------------------------
// SC-PML: 1/s(l), s = 1 - i sigma(l)/(omega eps0'),
// sigma(l) = sigma_max (l/T)^m, sigma_max = -(m+1) lnR / (2 eta0 T)
// (pml.jl:2-10); l at cell centres for the forward factor (:15-19) and at
// cell edges for the backward factor (:21-25); 1 outside the PML.  The
// reference stores the inverse 1/s in S_create (pml.jl:47-53).
// ---------------------------------------------------------------------------
__device__ inline double2 inv_sfactor(int i /*1-based*/, int Nw, int npml, double dw, bool fwd,
                                      double omega, double eps0n, double eta0) {
    if (npml == 0) return make_double2(1.0, 0.0);
    double T = npml * dw;
    double l;
    if (fwd) {
        if (i <= npml) l = dw * (npml - i + 0.5);
        else if (i > Nw - npml) l = dw * (i - (Nw - npml) - 0.5);
        else return make_double2(1.0, 0.0);
    } else {
        if (i <= npml) l = dw * (npml - i + 1);
        else if (i > Nw - npml) l = dw * (i - (Nw - npml) - 1);
        else return make_double2(1.0, 0.0);
    }
    double smax = -(PML_M + 1.0) * PML_LNR / (2.0 * eta0 * T);
    double sig = smax * pow(l / T, PML_M);
    double2 s = make_double2(1.0, -sig / (omega * eps0n));
    return crcp(s);
}

__global__ void k_sfactor(double2* sxf, double2* sxb, double2* syf, double2* syb, GridDesc g,
                          double omega, double eps0n, double eta0) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < g.Nx) {
        sxf[t] = inv_sfactor(t + 1, g.Nx, g.npx, g.dx, true, omega, eps0n, eta0);
        sxb[t] = inv_sfactor(t + 1, g.Nx, g.npx, g.dx, false, omega, eps0n, eta0);
    } else if (t < g.Nx + g.Ny) {
        int j = t - g.Nx;
        syf[j] = inv_sfactor(j + 1, g.Ny, g.npy, g.dy, true, omega, eps0n, eta0);
        syb[j] = inv_sfactor(j + 1, g.Ny, g.npy, g.dy, false, omega, eps0n, eta0);
    }
}

void launch_sfactor(hipStream_t s, double2* sxf, double2* sxb, double2* syf, double2* syb,
                    const GridDesc& g, double omega, double eps0n, double eta0) {
    int tot = g.Nx + g.Ny;
    k_sfactor<<<(tot + 255) / 256, 256, 0, s>>>(sxf, sxb, syf, syb, g, omega, eps0n, eta0);
}

// ---------------------------------------------------------------------------
// Coefficients of A' (scaled operator).
// TM (driven.jl:35), scale mu0' dx^2:
//   cE = sxf_i sxb_{i+1}, cW = sxf_i sxb_i, cN = r syf_j syb_{j+1},
//   cS = r syf_j syb_j (r = dx^2/dy^2), cC = -(cE+cW+cN+cS) + kdiag*eps_ij
// TE (driven.jl:45), scale eps0' dx^2, tx_ij = 2/(eps_ij+eps_{i-1,j})
// (= eps0'/grid_average(eps0' eps, x), grid.jl:157-162, periodic):
//   cE = sxf_i tx_{i+1,j} sxb_{i+1}, cW = sxf_i tx_ij sxb_i, (y analogous),
//   cC = -(sum) + kdiag.   Indices wrap periodically (grid.jl:138-150).
// ---------------------------------------------------------------------------
__global__ void k_coeffs(double2* coef, const double2* eps, const double2* sxf,
                         const double2* sxb, const double2* syf, const double2* syb, int Nx,
                         int Ny, double ry, int pol, double kdiag, int form) {
    long N = (long)Nx * Ny;
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    int i = (int)(k % Nx), j = (int)(k / Nx);
    int ip = (i + 1 == Nx) ? 0 : i + 1, im = (i == 0) ? Nx - 1 : i - 1;
    int jp = (j + 1 == Ny) ? 0 : j + 1, jm = (j == 0) ? Ny - 1 : j - 1;
    double2 cE, cW, cN, cS, cC;
    double2 ryc = make_double2(ry, 0.0);
    if (pol == 1 && form == 1) {
        // sideband operator of the modulated solve, Sxb dxb Sxf dxf + Syb dyb
        // Syf dyf (modulation.jl:83,88): sxb_i [sxf_i (e_E - e_C) - sxf_{i-1} (e_C - e_W)]
        cE = cmul(sxb[i], sxf[i]);
        cW = cmul(sxb[i], sxf[im]);
        cN = cmul(ryc, cmul(syb[j], syf[j]));
        cS = cmul(ryc, cmul(syb[j], syf[jm]));
        cC = cadd(cneg(cadd(cadd(cE, cW), cadd(cN, cS))), cscale(eps[k], kdiag));
    } else if (pol == 1) {
        cE = cmul(sxf[i], sxb[ip]);
        cW = cmul(sxf[i], sxb[i]);
        cN = cmul(ryc, cmul(syf[j], syb[jp]));
        cS = cmul(ryc, cmul(syf[j], syb[j]));
        cC = cadd(cneg(cadd(cadd(cE, cW), cadd(cN, cS))), cscale(eps[k], kdiag));
    } else {
        double2 e0 = eps[k];
        double2 two = make_double2(2.0, 0.0);
        double2 tx_i = cdiv(two, cadd(e0, eps[im + (long)j * Nx]));
        double2 tx_ip = cdiv(two, cadd(eps[ip + (long)j * Nx], e0));
        double2 ty_j = cdiv(two, cadd(e0, eps[i + (long)jm * Nx]));
        double2 ty_jp = cdiv(two, cadd(eps[i + (long)jp * Nx], e0));
        cE = cmul(sxf[i], cmul(tx_ip, sxb[ip]));
        cW = cmul(sxf[i], cmul(tx_i, sxb[i]));
        cN = cmul(ryc, cmul(syf[j], cmul(ty_jp, syb[jp])));
        cS = cmul(ryc, cmul(syf[j], cmul(ty_j, syb[j])));
        cC = cadd(cneg(cadd(cadd(cE, cW), cadd(cN, cS))), make_double2(kdiag, 0.0));
    }
    coef[CW * N + k] = cW;
    coef[CE * N + k] = cE;
    coef[CS * N + k] = cS;
    coef[CN * N + k] = cN;
    coef[CC * N + k] = cC;
}

void launch_coeffs(hipStream_t s, double2* coef, const double2* eps, const double2* sxf,
                   const double2* sxb, const double2* syf, const double2* syb,
                   const GridDesc& g, int pol, double kdiag, int form) {
    long N = (long)g.Nx * g.Ny;
    double ry = (g.dx * g.dx) / (g.dy * g.dy);
    k_coeffs<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(coef, eps, sxf, sxb, syf, syb, g.Nx,
                                                         g.Ny, ry, pol, kdiag, form);
}

// Sideband coupling of the modulated operator (modulation.jl:95-99):
// block (j, j+1) = 0.5 wn_j^2 conj(T_deps), block (j+1, j) = 0.5 wn_{j+1}^2 T_deps,
// so row block j gets c_j (conj(d) x_{j+1} + d x_{j-1}), c_j = scale 0.5 wn_j^2 eps0'.
__global__ void k_mf_couple(double2* y, const double2* x, const double2* d, const double* c,
                            int nf, long N, double sign) {
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double2 dk = d[k];
    if (dk.x == 0.0 && dk.y == 0.0) return;
    for (int j = 0; j < nf; ++j) {
        double2 acc = make_double2(0.0, 0.0);
        if (j + 1 < nf) acc = cadd(acc, cmul(make_double2(dk.x, -dk.y), x[(long)(j + 1) * N + k]));
        if (j > 0) acc = cadd(acc, cmul(dk, x[(long)(j - 1) * N + k]));
        const long o = (long)j * N + k;
        y[o] = cadd(y[o], cscale(acc, sign * c[j]));
    }
}

void launch_mf_couple(hipStream_t s, double2* y, const double2* x, const double2* d,
                      const double* c, int nf, long N, double sign) {
    k_mf_couple<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(y, x, d, c, nf, N, sign);
}

// ---------------------------------------------------------------------------
// y = A' x  (or A'^T x); if b != nullptr: y = b - A' x.
// A'^T x at (i,j) = cE_{i-1} x_{i-1} + cW_{i+1} x_{i+1} + cN_{j-1} x_{j-1}
//                 + cS_{j+1} x_{j+1} + cC x.
// ---------------------------------------------------------------------------
__global__ void k_apply(double2* y, const double2* x, const double2* coef, int Nx, int Ny,
                        int transpose, const double2* b) {
    long N = (long)Nx * Ny;
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    int i = (int)(k % Nx), j = (int)(k / Nx);
    long kw = (i == 0) ? k + Nx - 1 : k - 1;
    long ke = (i + 1 == Nx) ? k - (Nx - 1) : k + 1;
    long ks = (j == 0) ? k + N - Nx : k - Nx;
    long kn = (j + 1 == Ny) ? k - (N - Nx) : k + Nx;
    double2 acc;
    if (!transpose) {
        acc = cmul(coef[CC * N + k], x[k]);
        acc = cadd(acc, cmul(coef[CW * N + k], x[kw]));
        acc = cadd(acc, cmul(coef[CE * N + k], x[ke]));
        acc = cadd(acc, cmul(coef[CS * N + k], x[ks]));
        acc = cadd(acc, cmul(coef[CN * N + k], x[kn]));
    } else {
        acc = cmul(coef[CC * N + k], x[k]);
        acc = cadd(acc, cmul(coef[CE * N + kw], x[kw]));
        acc = cadd(acc, cmul(coef[CW * N + ke], x[ke]));
        acc = cadd(acc, cmul(coef[CN * N + ks], x[ks]));
        acc = cadd(acc, cmul(coef[CS * N + kn], x[kn]));
    }
    y[k] = b ? csub(b[k], acc) : acc;
}

void launch_apply(hipStream_t s, double2* y, const double2* x, const double2* coef, int Nx,
                  int Ny, int transpose, const double2* b) {
    long N = (long)Nx * Ny;
    k_apply<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(y, x, coef, Nx, Ny, transpose, b);
}

// ---------------------------------------------------------------------------
// Matrix-free form of the same operator (driven TM/TE, form 0, and the TM
// sideband form 1): the stencil weights are formed in registers from the 1-D
// inverse stretch vectors (a few KB, cache-resident) and eps, so the kernel
// streams x, eps, b and y only (64 B per cell) instead of five coefficient
// planes.  A workgroup owns a 64 x 16 tile: x (and eps for TE, whose weights
// need the neighbours' averages) is staged with a one-cell periodic halo in
// LDS, each of the 256 lanes computes one column of four cells, and the
// loads and stores of a tile row are contiguous.  The weights are the ones
// k_coeffs writes, in the same operation order:
//   A'x  at (i,j) = cW x_W + cE x_E + cS x_S + cN x_N + cC x_C
//   A'^T x at (i,j) = cE(i-1) x_W + cW(i+1) x_E + cN(j-1) x_S + cS(j+1) x_N + cC x_C
// ---------------------------------------------------------------------------
constexpr int ST_TX = 64, ST_TY = 16;

namespace {
struct StW { double2 w, e, s, n, c; };

// weights of cell (i, j); ev(di, dj) returns eps at (i+di, j+dj) (TE only)
template <int POL, int FORM, typename EV>
__device__ inline StW st_weights(int i, int j, int Nx, int Ny, const double2* sxf,
                                 const double2* sxb, const double2* syf, const double2* syb,
                                 double ry, double kdiag, EV ev) {
    const int ip = (i + 1 == Nx) ? 0 : i + 1, im = (i == 0) ? Nx - 1 : i - 1;
    const int jp = (j + 1 == Ny) ? 0 : j + 1, jm = (j == 0) ? Ny - 1 : j - 1;
    const double2 ryc = make_double2(ry, 0.0);
    StW W;
    if (POL == 1 && FORM == 1) {
        W.e = cmul(sxb[i], sxf[i]);
        W.w = cmul(sxb[i], sxf[im]);
        W.n = cmul(ryc, cmul(syb[j], syf[j]));
        W.s = cmul(ryc, cmul(syb[j], syf[jm]));
        W.c = cadd(cneg(cadd(cadd(W.e, W.w), cadd(W.n, W.s))), cscale(ev(0, 0), kdiag));
    } else if (POL == 1) {
        W.e = cmul(sxf[i], sxb[ip]);
        W.w = cmul(sxf[i], sxb[i]);
        W.n = cmul(ryc, cmul(syf[j], syb[jp]));
        W.s = cmul(ryc, cmul(syf[j], syb[j]));
        W.c = cadd(cneg(cadd(cadd(W.e, W.w), cadd(W.n, W.s))), cscale(ev(0, 0), kdiag));
    } else {
        const double2 e0 = ev(0, 0);
        const double2 two = make_double2(2.0, 0.0);
        const double2 tx_i = cdiv(two, cadd(e0, ev(-1, 0)));
        const double2 tx_ip = cdiv(two, cadd(ev(1, 0), e0));
        const double2 ty_j = cdiv(two, cadd(e0, ev(0, -1)));
        const double2 ty_jp = cdiv(two, cadd(ev(0, 1), e0));
        W.e = cmul(sxf[i], cmul(tx_ip, sxb[ip]));
        W.w = cmul(sxf[i], cmul(tx_i, sxb[i]));
        W.n = cmul(ryc, cmul(syf[j], cmul(ty_jp, syb[jp])));
        W.s = cmul(ryc, cmul(syf[j], cmul(ty_j, syb[j])));
        W.c = cadd(cneg(cadd(cadd(W.e, W.w), cadd(W.n, W.s))), make_double2(kdiag, 0.0));
    }
    return W;
}
}  // namespace

template <int POL, int FORM, bool TRANS>
__global__ __launch_bounds__(256) void k_stencil(double2* __restrict__ y, const double2* __restrict__ x,
                                                 const double2* __restrict__ eps,
                                                 const double2* sxf, const double2* sxb,
                                                 const double2* syf, const double2* syb, int Nx,
                                                 int Ny, double ry, double kdiag,
                                                 const double2* __restrict__ b, long rs,
                                                 const int* ridx) {
    // several right-hand sides: grid z -> source ridx[z], vectors at r * rs
    if (ridx) {
        const long ro = (long)ridx[blockIdx.z] * rs;
        y += ro;
        x += ro;
        if (b) b += ro;
    }
    // TE weights (and every transposed weight) need eps one or two cells away
    constexpr int EH = (POL == 2) ? (TRANS ? 2 : 1) : 0;
    __shared__ double2 xs[ST_TY + 2][ST_TX + 2];
    __shared__ double2 es[(POL == 2) ? ST_TY + 2 * EH : 1][(POL == 2) ? ST_TX + 2 * EH : 1];
    const int i0 = blockIdx.x * ST_TX, j0 = blockIdx.y * ST_TY;
    const int t = threadIdx.x;
    auto wrapx = [&](int v) { v %= Nx; return v < 0 ? v + Nx : v; };
    auto wrapy = [&](int v) { v %= Ny; return v < 0 ? v + Ny : v; };
    for (int q = t; q < (ST_TY + 2) * (ST_TX + 2); q += 256) {
        const int r = q / (ST_TX + 2), c = q % (ST_TX + 2);
        const int gi = wrapx(i0 - 1 + c), gj = wrapy(j0 - 1 + r);
        xs[r][c] = x[(long)gj * Nx + gi];
    }
    if (POL == 2) {
        constexpr int EW = ST_TX + 2 * EH, EHt = ST_TY + 2 * EH;
        for (int q = t; q < EHt * EW; q += 256) {
            const int r = q / EW, c = q % EW;
            const int gi = wrapx(i0 - EH + c), gj = wrapy(j0 - EH + r);
            es[r][c] = eps[(long)gj * Nx + gi];
        }
    }
    __syncthreads();
    const int lx = t & (ST_TX - 1), ly0 = t >> 6;
    const int i = i0 + lx;
    if (i >= Nx) return;
#pragma unroll
    for (int q = 0; q < ST_TY / 4; ++q) {
        const int ly = ly0 + 4 * q, j = j0 + ly;
        if (j >= Ny) break;
        const long k = (long)j * Nx + i;
        // eps at (i + di, j + dj) relative to the weight's cell (ci, cj)
        auto weights = [&](int ci, int cj, int oi, int oj) {
            auto ev = [&](int di, int dj) -> double2 {
                if (POL == 2) return es[ly + EH + oj + dj][lx + EH + oi + di];
                return eps[k];
            };
            return st_weights<POL, FORM>(ci, cj, Nx, Ny, sxf, sxb, syf, syb, ry, kdiag, ev);
        };
        const double2 xc = xs[ly + 1][lx + 1];
        const double2 xw = xs[ly + 1][lx], xe = xs[ly + 1][lx + 2];
        const double2 xsn = xs[ly][lx + 1], xn = xs[ly + 2][lx + 1];
        double2 acc;
        const StW W = weights(i, j, 0, 0);
        if (!TRANS) {
            acc = cmul(W.c, xc);
            acc = cadd(acc, cmul(W.w, xw));
            acc = cadd(acc, cmul(W.e, xe));
            acc = cadd(acc, cmul(W.s, xsn));
            acc = cadd(acc, cmul(W.n, xn));
        } else {
            const int im = (i == 0) ? Nx - 1 : i - 1, ip = (i + 1 == Nx) ? 0 : i + 1;
            const int jm = (j == 0) ? Ny - 1 : j - 1, jp = (j + 1 == Ny) ? 0 : j + 1;
            acc = cmul(W.c, xc);
            acc = cadd(acc, cmul(weights(im, j, -1, 0).e, xw));
            acc = cadd(acc, cmul(weights(ip, j, 1, 0).w, xe));
            acc = cadd(acc, cmul(weights(i, jm, 0, -1).n, xsn));
            acc = cadd(acc, cmul(weights(i, jp, 0, 1).s, xn));
        }
        y[k] = b ? csub(b[k], acc) : acc;
    }
}

void launch_stencil(hipStream_t s, double2* y, const double2* x, const StencilOp& op,
                    int transpose, const double2* b, long rs, const int* ridx, int na) {
    const dim3 grid((op.Nx + ST_TX - 1) / ST_TX, (op.Ny + ST_TY - 1) / ST_TY, ridx ? na : 1);
#define ST_LAUNCH(P, F, T)                                                                     \
    k_stencil<P, F, T><<<grid, 256, 0, s>>>(y, x, op.eps, op.sxf, op.sxb, op.syf, op.syb, op.Nx, \
                                            op.Ny, op.ry, op.kdiag, b, rs, ridx)
    if (op.form == 1) {
        if (transpose) ST_LAUNCH(1, 1, true);
        else ST_LAUNCH(1, 1, false);
    } else if (op.pol == 1) {
        if (transpose) ST_LAUNCH(1, 0, true);
        else ST_LAUNCH(1, 0, false);
    } else {
        if (transpose) ST_LAUNCH(2, 0, true);
        else ST_LAUNCH(2, 0, false);
    }
#undef ST_LAUNCH
}

// y = a * x
__global__ void k_scale_copy(double2* y, const double2* x, double2 a, long n) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; k < n; k += (long)gridDim.x * blockDim.x) y[k] = cmul(a, x[k]);
}
void launch_scale_copy(hipStream_t s, double2* y, const double2* x, double2 a, long n) {
    unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 4096);
    k_scale_copy<<<nb, 256, 0, s>>>(y, x, a, n);
}

// y = 0
__global__ void k_zero(double2* y, long n) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; k < n; k += (long)gridDim.x * blockDim.x) y[k] = make_double2(0.0, 0.0);
}
void launch_zero(hipStream_t s, double2* y, long n) {
    unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 4096);
    k_zero<<<nb, 256, 0, s>>>(y, n);
}

// y = a * eps .* x
__global__ void k_mul_eps(double2* y, const double2* x, const double2* eps, double2 a, long n) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; k < n; k += (long)gridDim.x * blockDim.x) y[k] = cmul(a, cmul(eps[k], x[k]));
}
void launch_mul_eps(hipStream_t s, double2* y, const double2* x, const double2* eps, double2 a,
                    long n) {
    unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 4096);
    k_mul_eps<<<nb, 256, 0, s>>>(y, x, eps, a, n);
}

// ---------------------------------------------------------------------------
// reductions (deterministic two-pass: per-block partials, then one block)
// ---------------------------------------------------------------------------
template <typename T> __device__ inline T wave_sum(T v);
template <> __device__ inline double wave_sum<double>(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// partial[blk*k + i] = sum_{e in blk} conj(V_i[e]) w[e]   (k <= 32)
constexpr int MD_MAXK = 32;
__global__ __launch_bounds__(256) void k_multidot(const double2* V, long ldv, int k,
                                                  const double2* w, long n, double2* partial) {
    __shared__ double2 red[4][MD_MAXK];
    double2 acc[MD_MAXK];
    for (int i = 0; i < MD_MAXK; ++i) acc[i] = make_double2(0, 0);
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        double2 we = w[e];
#pragma unroll
        for (int i = 0; i < MD_MAXK; ++i)
            if (i < k) acc[i] = cadd(acc[i], cconjmul(V[i * ldv + e], we));
    }
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < MD_MAXK; ++i) {
        if (i < k) {
            double re = wave_sum(acc[i].x), im = wave_sum(acc[i].y);
            if (lane == 0) red[wv][i] = make_double2(re, im);
        }
    }
    __syncthreads();
    if (threadIdx.x < k) {
        double2 s = red[0][threadIdx.x];
        for (int q = 1; q < 4; ++q) s = cadd(s, red[q][threadIdx.x]);
        partial[blockIdx.x * (long)k + threadIdx.x] = s;
    }
}

void launch_multidot(hipStream_t s, const double2* V, long ldv, int k, const double2* w, long n,
                     double2* partial, int nblk) {
    k_multidot<<<nblk, 256, 0, s>>>(V, ldv, k, w, n, partial);
}

// Final sums of per-block partials, one workgroup of 1024 threads: thread t
// adds column t % 32 over the rows t / 32, t / 32 + 32, ... (a fixed order, so
// results are reproducible), then the 32 row groups of a column are added in
// LDS.  (One thread per column walking all rows waited out one load latency
// per row: 77 us for 512 rows.)
constexpr int RP_COLS = 32, RP_GROUPS = 32;
__device__ inline void reduce_columns(const double2* partial, int nblk, int k, double2* out) {
    __shared__ double2 acc[RP_GROUPS][RP_COLS];
    const int c = threadIdx.x % RP_COLS, grp = threadIdx.x / RP_COLS;
    for (int c0 = 0; c0 < k; c0 += RP_COLS) {
        const int i = c0 + c;
        double2 s = make_double2(0, 0);
        if (i < k)
            for (int b = grp; b < nblk; b += RP_GROUPS) s = cadd(s, partial[b * (long)k + i]);
        acc[grp][c] = s;
        __syncthreads();
        if (grp == 0 && i < k) {
            double2 t = acc[0][c];
            for (int g2 = 1; g2 < RP_GROUPS; ++g2) t = cadd(t, acc[g2][c]);
            out[i] = t;
        }
        __syncthreads();
    }
}
__global__ __launch_bounds__(1024) void k_reduce_partials(const double2* partial, int nblk, int k,
                                                          double2* out) {
    reduce_columns(partial, nblk, k, out);
}
void launch_reduce_partials(hipStream_t s, const double2* partial, int nblk, int k,
                            double2* out) {
    k_reduce_partials<<<1, RP_COLS * RP_GROUPS, 0, s>>>(partial, nblk, k, out);
}

// w += sign * sum_i h_i V_i
__global__ void k_gemv_update(double2* w, const double2* V, long ldv, int k, const double2* h,
                              double sign, long n) {
    __shared__ double2 hs[64];
    if (threadIdx.x < k) hs[threadIdx.x] = cscale(h[threadIdx.x], sign);
    __syncthreads();
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        double2 acc = w[e];
        for (int i = 0; i < k; ++i) acc = cadd(acc, cmul(hs[i], V[i * ldv + e]));
        w[e] = acc;
    }
}
void launch_gemv_update(hipStream_t s, double2* w, const double2* V, long ldv, int k,
                        const double2* h, double sign, long n) {
    unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 2048);
    k_gemv_update<<<nb, 256, 0, s>>>(w, V, ldv, k, h, sign, n);
}

__global__ __launch_bounds__(256) void k_norm2(const double2* x, long n, double* partial) {
    __shared__ double red[4];
    double acc = 0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x)
        acc += cnorm2(x[e]);
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
void launch_norm2(hipStream_t s, const double2* x, long n, double* partial, int nblk) {
    k_norm2<<<nblk, 256, 0, s>>>(x, n, partial);
}
// sum of nblk partials: 256 threads, fixed strided order, then a fixed tree
__device__ inline void reduce_real(const double* partial, int nblk, double* out) {
    __shared__ double acc[256];
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[b];
    acc[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) acc[threadIdx.x] += acc[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = acc[0];
}
__global__ __launch_bounds__(256) void k_reduce_real(const double* partial, int nblk, double* out) {
    reduce_real(partial, nblk, out);
}
void launch_reduce_real(hipStream_t s, const double* partial, int nblk, double* out) {
    k_reduce_real<<<1, 256, 0, s>>>(partial, nblk, out);
}

// ---------------------------------------------------------------------------
// field recovery.  out has 3 planes of N: [u, c1, c2].
// driven TM (driven.jl:40-41): Hx = -w * syb_j (u_ij - u_i,j-1)/dy,
//                              Hy =  w * sxb_i (u_ij - u_i-1,j)/dx,  w = 1/(i omega mu0')
// eigen TM (eigen.jl:90-91): forward differences syf_j (u_i,j+1 - u_ij)/dy.
// driven TE (driven.jl:50-51), w = 1/(i omega):
//   Ex = w * Teyi_ij * syb_j (u_ij - u_i,j-1)/dy,
//   Ey = -w * Texi_ij * sxb_i (u_ij - u_i-1,j)/dx,  Texi = 1/(eps0' avg_x eps)
// eigen TE (eigen.jl:108-109): Ex uses Texi, Ey uses Teyi.
// eigen_mode: 0 driven, 1 eigen.  The eps0' of TE is folded into w by the host.
// ---------------------------------------------------------------------------
__global__ void k_fields(double2* out, const double2* x, const double2* eps, const double2* sxf,
                         const double2* sxb, const double2* syf, const double2* syb, int Nx,
                         int Ny, double dx, double dy, int pol, double2 w, int eigen_mode) {
    long N = (long)Nx * Ny;
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    int i = (int)(k % Nx), j = (int)(k / Nx);
    long kw = (i == 0) ? k + Nx - 1 : k - 1;
    long ke = (i + 1 == Nx) ? k - (Nx - 1) : k + 1;
    long ks = (j == 0) ? k + N - Nx : k - Nx;
    long kn = (j + 1 == Ny) ? k - (N - Nx) : k + Nx;
    double2 u = x[k];
    out[k] = u;
    if (pol == 1) {
        double2 dyv, dxv;
        if (!eigen_mode) {
            dyv = cscale(cmul(syb[j], csub(u, x[ks])), 1.0 / dy);
            dxv = cscale(cmul(sxb[i], csub(u, x[kw])), 1.0 / dx);
        } else {
            dyv = cscale(cmul(syf[j], csub(x[kn], u)), 1.0 / dy);
            dxv = cscale(cmul(sxf[i], csub(x[ke], u)), 1.0 / dx);
        }
        out[N + k] = cneg(cmul(w, dyv));
        out[2 * N + k] = cmul(w, dxv);
    } else {
        double2 two = make_double2(2.0, 0.0);
        // 1/avg = 2/(eps_ij + eps_neighbour); eps0' is inside w
        double2 tx = cdiv(two, cadd(eps[k], eps[kw]));
        double2 ty = cdiv(two, cadd(eps[k], eps[ks]));
        double2 dyv = cscale(cmul(syb[j], csub(u, x[ks])), 1.0 / dy);
        double2 dxv = cscale(cmul(sxb[i], csub(u, x[kw])), 1.0 / dx);
        double2 tex = eigen_mode ? tx : ty;
        double2 tey = eigen_mode ? ty : tx;
        out[N + k] = cmul(w, cmul(tex, dyv));
        out[2 * N + k] = cneg(cmul(w, cmul(tey, dxv)));
    }
}

void launch_fields(hipStream_t s, double2* out, const double2* x, const double2* eps,
                   const double2* sxf, const double2* sxb, const double2* syf,
                   const double2* syb, const GridDesc& g, int pol, double2 w,
                   int eigen_mode) {
    long N = (long)g.Nx * g.Ny;
    k_fields<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(out, x, eps, sxf, sxb, syf, syb, g.Nx,
                                                         g.Ny, g.dx, g.dy, pol, w, eigen_mode);
}

// ---------------------------------------------------------------------------
// Poynting vector and flux through planes (src/flux.jl:15-66).  Fields are
// the 3 x N layout of fdfd_solve (Ez, Hx, Hy or Hz, Ex, Ey), x fastest.
//   TM: Sx = -1/2 Re(avg_x(Ez) conj(Hy)),  Sy = 1/2 Re(avg_y(Ez) conj(Hx))
//   TE: Sx =  1/2 Re(Ey conj(avg_x(Hz))),  Sy = -1/2 Re(Ex conj(avg_y(Hz)))
// avg = grid_average (periodic, the cell and its lower neighbour).
// ---------------------------------------------------------------------------
__global__ void k_poynting(double* S, const double2* f, int Nx, int Ny, int pol) {
    const long N = (long)Nx * Ny;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const long kw = (i == 0) ? k + Nx - 1 : k - 1;
    const long ks = (j == 0) ? k + N - Nx : k - Nx;
    const double2 u = f[k];
    const double2 ax = cscale(cadd(u, f[kw]), 0.5), ay = cscale(cadd(u, f[ks]), 0.5);
    const double2 a = f[N + k], b = f[2 * N + k];  // Hx, Hy (TM) or Ex, Ey (TE)
    if (pol == 1) {
        S[k] = -0.5 * cmulc(ax, b).x;      // Re(ax conj(Hy))
        S[N + k] = 0.5 * cmulc(ay, a).x;   // Re(ay conj(Hx))
    } else {
        S[k] = 0.5 * cmulc(b, ax).x;       // Re(Ey conj(ax))
        S[N + k] = -0.5 * cmulc(a, ay).x;  // Re(Ex conj(ay))
    }
}

void launch_poynting(hipStream_t s, double* S, const double2* f, int Nx, int Ny, int pol) {
    const long N = (long)Nx * Ny;
    k_poynting<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(S, f, Nx, Ny, pol);
}

// One workgroup per plane {normal, index, lo, hi} (0-based, lo..hi inclusive
// along the plane, index + 1 < the normal extent): the normal component with
// the first field averaged over the plane's line and the next one (the
// reference's view xind:xind+1, flux.jl:42-44), summed, times the cell size
// along the plane.
//   x normal, TM: -1/2 Re(avg Ez conj(Hy))   TE: 1/2 Re(Ey conj(avg Hz))
//   y normal, TM:  1/2 Re(avg Ez conj(Hx))   TE: -1/2 Re(Ex conj(avg Hz))
__global__ __launch_bounds__(256) void k_flux(double* out, const double2* f, const int4* planes,
                                              int Nx, int Ny, int pol, double dx, double dy) {
    const long N = (long)Nx * Ny;
    const int4 P = planes[blockIdx.x];
    const bool xn = P.x == 1;
    double acc = 0;
    for (int t = P.z + (int)threadIdx.x; t <= P.w; t += 256) {
        const long k = xn ? (long)t * Nx + P.y : (long)P.y * Nx + t;
        const long k1 = xn ? k + 1 : k + Nx;
        const double2 u = cscale(cadd(f[k], f[k1]), 0.5);
        double v;
        if (xn) v = (pol == 1) ? -0.5 * cmulc(u, f[2 * N + k]).x : 0.5 * cmulc(f[2 * N + k], u).x;
        else v = (pol == 1) ? 0.5 * cmulc(u, f[N + k]).x : -0.5 * cmulc(f[N + k], u).x;
        acc += v;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (part[0] + part[1] + part[2] + part[3]) * (xn ? dy : dx);
}

void launch_flux(hipStream_t s, double* out, const double2* f, const int4* planes, int nplanes,
                 int Nx, int Ny, int pol, double dx, double dy) {
    k_flux<<<nplanes, 256, 0, s>>>(out, f, planes, Nx, Ny, pol, dx, dy);
}

// ---------------------------------------------------------------------------
// Full-vector mode operator of a waveguide cross-section (eigen.jl:31-47),
// unknowns (Hx, Hy), periodic differences, no PML:
//   A_xx = k0^2 ey + ey dyf tz dyb + dxb dxf
//   A_xy = -ey dyf tz dxb + dxb dyf
//   A_yx = -ex dxf tz dyb + dyb dxf
//   A_yy = k0^2 ex + ex dxf tz dxb + dyb dyf
// with relative permittivity averages ex = avg_x eps, ey = avg_y eps,
// tz = 1 / avg_y(avg_x eps) (grid_average, grid.jl:157-162; eps0 cancels).
// Everything is scaled by s = dx^2 so the factors see O(1) entries.
// ---------------------------------------------------------------------------
namespace {
__device__ inline double2 vm_eps(const double2* e, int i, int j, int Nx, int Ny) {
    i = (i + Nx) % Nx;
    j = (j + Ny) % Ny;
    return e[(long)j * Nx + i];
}
__device__ inline double2 vm_ex(const double2* e, int i, int j, int Nx, int Ny) {
    return cscale(cadd(vm_eps(e, i, j, Nx, Ny), vm_eps(e, i - 1, j, Nx, Ny)), 0.5);
}
__device__ inline double2 vm_ey(const double2* e, int i, int j, int Nx, int Ny) {
    return cscale(cadd(vm_eps(e, i, j, Nx, Ny), vm_eps(e, i, j - 1, Nx, Ny)), 0.5);
}
__device__ inline double2 vm_tz(const double2* e, int i, int j, int Nx, int Ny) {
    const double2 a = cscale(cadd(vm_ex(e, i, j, Nx, Ny), vm_ex(e, i, j - 1, Nx, Ny)), 0.5);
    return crcp(a);
}
}  // namespace

// 5-point planes of s (A_bb - sigma I), block b = 0 (xx) or 1 (yy)
__global__ void k_vmode_coeffs(double2* coef, const double2* eps, int Nx, int Ny, double idx2,
                               double idy2, double k0sq, double2 sig, double s, int block) {
    const long N = (long)Nx * Ny;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    double2 cE, cW, cN, cS, diag;
    if (block == 0) {
        const double2 ey = vm_ey(eps, i, j, Nx, Ny);
        cE = make_double2(idx2, 0);
        cW = cE;
        cN = cscale(cmul(ey, vm_tz(eps, i, j + 1, Nx, Ny)), idy2);
        cS = cscale(cmul(ey, vm_tz(eps, i, j, Nx, Ny)), idy2);
        diag = cscale(ey, k0sq);
    } else {
        const double2 ex = vm_ex(eps, i, j, Nx, Ny);
        cE = cscale(cmul(ex, vm_tz(eps, i + 1, j, Nx, Ny)), idx2);
        cW = cscale(cmul(ex, vm_tz(eps, i, j, Nx, Ny)), idx2);
        cN = make_double2(idy2, 0);
        cS = cN;
        diag = cscale(ex, k0sq);
    }
    const double2 cC = csub(csub(diag, cadd(cadd(cE, cW), cadd(cN, cS))), sig);
    coef[CW * N + k] = cscale(cW, s);
    coef[CE * N + k] = cscale(cE, s);
    coef[CS * N + k] = cscale(cS, s);
    coef[CN * N + k] = cscale(cN, s);
    coef[CC * N + k] = cscale(cC, s);
}

void launch_vmode_coeffs(hipStream_t st, double2* coef, const double2* eps, const GridDesc& g,
                         double k0sq, double2 sig, double s, int block) {
    const long N = (long)g.Nx * g.Ny;
    k_vmode_coeffs<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(
        coef, eps, g.Nx, g.Ny, 1.0 / (g.dx * g.dx), 1.0 / (g.dy * g.dy), k0sq, sig, s, block);
}

// y[0:N] += c s A_xy x[N:2N],  y[N:2N] += c s A_yx x[0:N]
__global__ void k_vmode_offdiag(double2* y, const double2* x, const double2* eps, int Nx, int Ny,
                                double idxdy, double cs) {
    const long N = (long)Nx * Ny;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const double2* hx = x;
    const double2* hy = x + N;
    auto at = [&](const double2* v, int a, int b) { return v[(long)((b + Ny) % Ny) * Nx + (a + Nx) % Nx]; };
    const double2 tz = vm_tz(eps, i, j, Nx, Ny);
    // A_xy hy
    {
        const double2 ey = vm_ey(eps, i, j, Nx, Ny);
        const double2 tzn = vm_tz(eps, i, j + 1, Nx, Ny);
        const double2 dn = csub(at(hy, i, j + 1), at(hy, i - 1, j + 1));
        const double2 d0 = csub(at(hy, i, j), at(hy, i - 1, j));
        const double2 t1 = cmul(ey, csub(cmul(tzn, dn), cmul(tz, d0)));
        const double2 v = cscale(csub(csub(dn, d0), t1), idxdy * cs);
        y[k] = cadd(y[k], v);
    }
    // A_yx hx
    {
        const double2 ex = vm_ex(eps, i, j, Nx, Ny);
        const double2 tze = vm_tz(eps, i + 1, j, Nx, Ny);
        const double2 de = csub(at(hx, i + 1, j), at(hx, i + 1, j - 1));
        const double2 d0 = csub(at(hx, i, j), at(hx, i, j - 1));
        const double2 t1 = cmul(ex, csub(cmul(tze, de), cmul(tz, d0)));
        const double2 v = cscale(csub(csub(de, d0), t1), idxdy * cs);
        y[N + k] = cadd(y[N + k], v);
    }
}

void launch_vmode_offdiag(hipStream_t st, double2* y, const double2* x, const double2* eps,
                          const GridDesc& g, double cs) {
    const long N = (long)g.Nx * g.Ny;
    k_vmode_offdiag<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(y, x, eps, g.Nx, g.Ny,
                                                                 1.0 / (g.dx * g.dy), cs);
}

// fields of a mode (eigen.jl:56-61); out = Ex, Ey, Ez, Hx, Hy, Hz (6 N),
// h = (Hx, Hy) (2 N).  Two passes: Hz, then E from Hz's neighbours.
//   Hz = (dxf Hx + dyf Hy) / (i b)
//   Ex = (dyb Hz + i b Hy) / (i w eps0 ex),  Ey = (-i b Hx - dxb Hz) / (i w eps0 ey)
//   Ez = (dxb Hy - dyb Hx) tz / (i w eps0)
__global__ void k_vmode_hz(double2* out, const double2* h, int Nx, int Ny, double idx, double idy,
                           double2 inv_ib) {
    const long N = (long)Nx * Ny;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const long ke = (i + 1 == Nx) ? k - (Nx - 1) : k + 1;
    const long kn = (j + 1 == Ny) ? k - (N - Nx) : k + Nx;
    const double2 dv = cadd(cscale(csub(h[ke], h[k]), idx), cscale(csub(h[N + kn], h[N + k]), idy));
    out[5 * N + k] = cmul(dv, inv_ib);
    out[3 * N + k] = h[k];
    out[4 * N + k] = h[N + k];
}

__global__ void k_vmode_e(double2* out, const double2* eps, int Nx, int Ny, double idx, double idy,
                          double2 ib, double2 inv_iweps0) {
    const long N = (long)Nx * Ny;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const long kw = (i == 0) ? k + Nx - 1 : k - 1;
    const long ks = (j == 0) ? k + N - Nx : k - Nx;
    const double2 hx = out[3 * N + k], hy = out[4 * N + k], hz = out[5 * N + k];
    const double2 ex = cdiv(cadd(cscale(csub(hz, out[5 * N + ks]), idy), cmul(ib, hy)),
                            vm_ex(eps, i, j, Nx, Ny));
    const double2 ey = cdiv(csub(cneg(cmul(ib, hx)), cscale(csub(hz, out[5 * N + kw]), idx)),
                            vm_ey(eps, i, j, Nx, Ny));
    const double2 curl = csub(cscale(csub(hy, out[4 * N + kw]), idx), cscale(csub(hx, out[3 * N + ks]), idy));
    out[k] = cmul(ex, inv_iweps0);
    out[N + k] = cmul(ey, inv_iweps0);
    out[2 * N + k] = cmul(cmul(curl, vm_tz(eps, i, j, Nx, Ny)), inv_iweps0);
}

void launch_vmode_fields(hipStream_t st, double2* out, const double2* h, const double2* eps,
                         const GridDesc& g, double2 beta, double omega, double eps0n) {
    const long N = (long)g.Nx * g.Ny;
    const unsigned nb = (unsigned)((N + 255) / 256);
    const double2 ib = make_double2(-beta.y, beta.x);           // i beta
    const double2 inv_ib = crcp(ib);
    const double2 inv_iweps0 = make_double2(0.0, -1.0 / (omega * eps0n));
    k_vmode_hz<<<nb, 256, 0, st>>>(out, h, g.Nx, g.Ny, 1.0 / g.dx, 1.0 / g.dy, inv_ib);
    k_vmode_e<<<nb, 256, 0, st>>>(out, eps, g.Nx, g.Ny, 1.0 / g.dx, 1.0 / g.dy, ib, inv_iweps0);
}

// grad[k] = Re(coef * lam[k] * e[k])
__global__ void k_grad(double* grad, const double2* e, const double2* lam, double2 coef,
                       long n) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k < n) grad[k] = cmul(coef, cmul(lam[k], e[k])).x;
}
void launch_grad(hipStream_t s, double* grad, const double2* e, const double2* lam,
                 double2 coef, long n) {
    k_grad<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(grad, e, lam, coef, n);
}

// Right-hand side of the adjoint solve of a TM driven solve from the
// cotangent (bar Ez, bar Hx, bar Hy) of its fields (3 x N, x fastest).  The H
// recovery Hx = -c Syb dyb Ez, Hy = c Sxb dxb Ez (driven.jl:40-41,
// c = 1/(i w mu0')) is linear in Ez, so the cotangent reaching Ez is
//   bar e = bar Ez + Mx^H bar Hx + My^H bar Hy,
//   Mx^H = -c dyf conj(Syb),  My^H = c dxf conj(Sxb)
// (conj(c) = -c and dyb^T = -dyf for the periodic differences).  The adjoint
// solve needs A^T lam = conj(bar e); out = scale * conj(bar e), the scaled
// operator's right-hand side.
__global__ void k_field_adjoint(double2* out, const double2* dfield, const double2* sxb,
                                const double2* syb, int Nx, int Ny, double rdx, double rdy,
                                double2 c, double scale) {
    const long N = (long)Nx * Ny;
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const long ke = (i + 1 == Nx) ? k - (Nx - 1) : k + 1;
    const long kn = (j + 1 == Ny) ? k - (N - Nx) : k + Nx;
    const int ie = (i + 1 == Nx) ? 0 : i + 1;
    const int jn = (j + 1 == Ny) ? 0 : j + 1;
    const double2* ez = dfield;
    const double2* hx = dfield + N;
    const double2* hy = dfield + 2 * N;
    // dyf(conj(syb) hx) and dxf(conj(sxb) hy) at k
    const double2 ty = csub(cconjmul(syb[jn], hx[kn]), cconjmul(syb[j], hx[k]));
    const double2 tx = csub(cconjmul(sxb[ie], hy[ke]), cconjmul(sxb[i], hy[k]));
    const double2 d = mk<double2>(tx.x * rdx - ty.x * rdy, tx.y * rdx - ty.y * rdy);
    const double2 e = cadd(ez[k], cmul(c, d));
    out[k] = mk<double2>(scale * e.x, -scale * e.y);
}
void launch_field_adjoint(hipStream_t s, double2* out, const double2* dfield, const double2* sxb,
                          const double2* syb, const GridDesc& g, double2 c, double scale) {
    const long N = (long)g.Nx * g.Ny;
    k_field_adjoint<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(out, dfield, sxb, syb, g.Nx, g.Ny,
                                                                1.0 / g.dx, 1.0 / g.dy, c, scale);
}

}  // namespace fdfd
