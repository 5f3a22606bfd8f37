// Batched line-block inversion kernels of libfdfd (gfx950).
//
// Each job is a dense n x n block (one grid line of a reduction level),
// inverted in place by blocked Gauss-Jordan elimination with partial
// pivoting.  Pivot rows are not swapped; they are tracked by index.  For a
// panel of nb columns K with original values V = A[:, K] and pivot rows p:
//   * k_panel    : GEPP pivot selection on the not-yet-pivoted rows of V
//                  (one workgroup, one row per thread in registers), giving
//                  the factors of the pivot block V[p, :] = L U;
//   * k_wpanel   : the same on a window of the first 64 unpivoted rows (the
//                  default; every pivot is then checked against all other
//                  rows by k_zbuild, threshold partial pivoting);
//   * k_zbuild   : Z = U^-1 L^-1 [A[p, C] | I_K]            (nb x n);
//   * k_trailing : A[i, :] = A'[i, :] - V[i, :] Z  (i not in p), A[p, :] = Z,
//                  a rank-nb complex update on MFMA (A' = A with K zeroed).
// With the window, two panels form a pair applied by one trailing pass: the
// second panel works on its columns as the first leaves them (formed by the
// second k_wpanel for its window and by the second k_zbuild for every row).
// After the last panel the stored Y holds the inverse X = D^-1 as
// X[a, b] = Y[prow[a], iprow[b]].
#include "fdfd_common.h"
#include "k_device.h"

#include <algorithm>
#include <type_traits>

namespace fdfd {



// ---------------------------------------------------------------------------
// Broadcast of a lane's value to the whole wave (lane index wave-uniform).
// ---------------------------------------------------------------------------
__device__ inline float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double bcast(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 64-bit DPP move within a 16-lane row (all 64 lanes active)
template <int CTRL> __device__ inline unsigned long long dpp64(unsigned long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffffULL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ inline unsigned long long readlane64(unsigned long long v, int l) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(v & 0xffffffffULL), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}
// wave-wide maximum of an unsigned 64-bit key: quad xor-1 / xor-2 and row
// rotations by 4 and 8 give every lane its row maximum, then the four rows
// are combined through scalar reads.  Result is wave-uniform.
__device__ inline unsigned long long wave_max_u64(unsigned long long k) {
    unsigned long long o;
    o = dpp64<0xB1>(k);  k = o > k ? o : k;   // quad_perm [1,0,3,2]
    o = dpp64<0x4E>(k);  k = o > k ? o : k;   // quad_perm [2,3,0,1]
    o = dpp64<0x124>(k); k = o > k ? o : k;   // row_ror:4
    o = dpp64<0x128>(k); k = o > k ? o : k;   // row_ror:8
    unsigned long long a = readlane64(k, 0), b = readlane64(k, 16);
    unsigned long long c = readlane64(k, 32), d = readlane64(k, 48);
    a = b > a ? b : a;
    c = d > c ? d : c;
    return c > a ? c : a;
}
// monotone key of a non-negative magnitude
// (sign bit cleared: -0 orders as +0; a NaN orders above +inf, so a NaN
// candidate is picked and flagged rather than hidden)
__device__ inline unsigned long long mag_key(double m) {
    return (unsigned long long)__double_as_longlong(m) & 0x7fffffffffffffffULL;
}
__device__ inline unsigned long long mag_key(float m) {
    return (unsigned long long)((unsigned)__float_as_int(m) & 0x7fffffffu);
}
// the same keys in the narrowest integer that holds them
__device__ inline unsigned long long wave_max_key(unsigned long long k) { return wave_max_u64(k); }
__device__ inline unsigned long long gepp_key(double m) { return mag_key(m); }
__device__ inline unsigned gepp_key(float m) { return (unsigned)__float_as_int(m) & 0x7fffffffu; }
__device__ inline unsigned wave_max_key(unsigned k) {
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0xB1, 0xF, 0xF, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x4E, 0xF, 0xF, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x124, 0xF, 0xF, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x128, 0xF, 0xF, false));
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)k, 0);
    const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)k, 16);
    const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)k, 32);
    const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)k, 48);
    return max(max(a, b), max(c, d));
}

// lane k of this lane's half-wave (lanes 0-31 or 32-63), k wave-uniform
template <typename C> __device__ inline C bcast_half(C v, int k) {
    // both halves' values are read (scalar reads, no divergence), then selected
    const bool hi = (threadIdx.x & 32) != 0;
    const auto x0 = bcast(v.x, k), x1 = bcast(v.x, 32 + k);
    const auto y0 = bcast(v.y, k), y1 = bcast(v.y, 32 + k);
    return mk<C>(hi ? x1 : x0, hi ? y1 : y0);
}

// Inverse factors of the pivot block V[piv, :] = L U (unit lower L, upper U,
// nb x nb in LDS) to the job's LU buffer: M = U^-1 L^-1 = V[piv, :]^-1 in the
// first PMAX x PMAX block, U^-1 in the second, both zero outside nb x nb.
// A half-wave solves for one column j of L^-1 (forward) and of U^-1
// (backward) at a time: lane i holds row i and each solved entry is
// broadcast with readlane.  M is then
// formed as the product U^-1 L^-1 through LDS (Lis, Uis: scratch).  Called by
// all nthr threads of the block (a multiple of 64) with Ls, Us complete.
template <typename C>
__device__ __attribute__((always_inline)) inline void inverse_product(C (*Lis)[PMAX + 1], C (*Uis)[PMAX + 1],
                                                                      int t, int nthr, C* LUg) {
    using R = decltype(C{}.x);
    const C zero = mk<C>(R(0), R(0));
    C* Mg = LUg;
    C* Ug = LUg + PMAX * PMAX;
    for (int e = t; e < PMAX * PMAX; e += nthr) {
        const int r = e / PMAX, c = e % PMAX;
        C sacc = zero;
#pragma unroll
        for (int k = 0; k < PMAX; ++k) sacc = cadd(sacc, cmul(Uis[r][k], Lis[k][c]));
        gstore(Mg + e, sacc);
        gstore(Ug + e, Uis[r][c]);
    }
}

template <typename C>
__device__ __attribute__((always_inline)) inline void store_pivot_inverses(C (*Ls)[PMAX + 1], C (*Us)[PMAX + 1],
                                            C (*Lis)[PMAX + 1], C (*Uis)[PMAX + 1], int nb,
                                            int t, int nthr, C* LUg) {
    using R = decltype(C{}.x);
    const C zero = mk<C>(R(0), R(0)), one = mk<C>(R(1), R(0));
    const int i = t & 31;
    const bool rok = i < nb;
    const C dinv = csel(rok, crcp(Us[i][i]), zero);
    for (int j0 = 2 * (t >> 6); j0 < PMAX; j0 += nthr / 32) {
        const int j = j0 + ((t >> 5) & 1);
        C x = csel(i == j, one, zero);  // L^-1 e_j
        C y = csel(i == j, one, zero);  // U^-1 e_j
        // rows and columns >= nb of Ls, Us are never used
        for (int k = 0; k < nb; ++k) {
            const C xk = bcast_half(x, k);
            x = csel(i > k && rok, csub(x, cmul(Ls[i][k], xk)), x);
            const int kb = nb - 1 - k;
            y = csel(i == kb, cmul(y, dinv), y);
            const C yk = bcast_half(y, kb);
            y = csel(i < kb, csub(y, cmul(Us[i][kb], yk)), y);
        }
        const bool ok = rok && j < nb;
        Lis[i][j] = csel(ok, x, zero);
        Uis[i][j] = csel(ok, y, zero);
    }
    __syncthreads();
    inverse_product(Lis, Uis, t, nthr, LUg);
}

// The same for the window panel, NTHR threads, with Ls holding the strictly
// lower part of L (zero diagonal) and both Ls and Us zero outside their
// leading nb x nb block, so that every substitution step runs
// unconditionally (steps past nb add exact zeros to the rows below nb).
// U = D (I + N) with N = D^-1 (U - D) strictly upper: the rows of U are
// scaled by 1 / U[i][i] in place first (Dv keeps the 1 / U[i][i]), so both
// substitutions are unit triangular, x <- x - N[:, k] x_k, and U^-1 e_j is
// (I + N)^-1 e_j times 1 / U[j][j].  Each wave carries whole complex
// columns: lane i holds the real part of row i, lane 32 + i its imaginary
// part, so a step is two readlanes (x_k) and two FMAs with the per-lane
// coefficients (re lanes: Re N, -Im N; im lanes: Im N, Re N) for each of
// its PMAX / NW columns of L^-1 and of U^-1 (j = w, w + NW, ..), all
// interleaved and fully unrolled (lane indices and LDS columns constant).
template <int NTHR, typename C>
__device__ __attribute__((always_inline)) inline void store_pivot_inverses_w(C (*Ls)[PMAX + 1], C (*Us)[PMAX + 1],
                                              C (*Lis)[PMAX + 1], C (*Uis)[PMAX + 1], decltype(C{}.x)* Dv,
                                              int nb, int t, C* LUg) {
    using R = decltype(C{}.x);
    constexpr int NW = NTHR / 64, NJ = PMAX / NW;
    static_assert(NJ * NW == PMAX, "columns per wave");
    const C zero = mk<C>(R(0), R(0));
    // U rows scaled by their inverse diagonal (values read before the barrier)
    {
        constexpr int EPT = PMAX * PMAX / NTHR;
        C nv[EPT];
        R dre = R(0), dim = R(0);
#pragma unroll
        for (int e_ = 0; e_ < EPT; ++e_) {
            const int e = t + NTHR * e_, r = e / PMAX, c = e % PMAX;
            const C d = (r < nb) ? crcp_scaled(Us[r][r]) : zero;
            nv[e_] = csel(c > r, cmul(d, Us[r][c]), zero);
            if (c == r) {
                dre = d.x;
                dim = d.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int e_ = 0; e_ < EPT; ++e_) {
            const int e = t + NTHR * e_, r = e / PMAX, c = e % PMAX;
            Us[r][c] = nv[e_];
            if (c == r) {
                Dv[2 * r] = dre;
                Dv[2 * r + 1] = dim;
            }
        }
        __syncthreads();
    }
    const int lane = t & 63, i = lane & 31;
    const bool him = lane >= 32;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    R x[NJ], y[NJ];  // this lane's part of L^-1 e_j, (I + N)^-1 e_j, j = w + NW q
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
        x[q] = (!him && i == w + NW * q) ? R(1) : R(0);
        y[q] = x[q];
    }
    static_for<0, PMAX>([&](auto ik) {
        constexpr int k = decltype(ik)::value, kb = PMAX - 1 - k;
        const C lv = Ls[i][k], uv = Us[i][kb];
        const R lp = him ? lv.y : lv.x, lq = him ? lv.x : -lv.y;
        const R up = him ? uv.y : uv.x, uq = him ? uv.x : -uv.y;
#pragma unroll
        for (int q = 0; q < NJ; ++q) {
            const R xr = bcast(x[q], k), xi = bcast(x[q], 32 + k);
            x[q] = fma(-lp, xr, fma(-lq, xi, x[q]));
            const R yr = bcast(y[q], kb), yi = bcast(y[q], 32 + kb);
            y[q] = fma(-up, yr, fma(-uq, yi, y[q]));
        }
    });
    R* Lr = (R*)&Lis[0][0];
    R* Ur = (R*)&Uis[0][0];
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
        const int j = w + NW * q;
        const bool ok = i < nb && j < nb;
        // U^-1 e_j = y / U[j][j]: the partner lane holds the other part
        const R yo = __shfl_xor(y[q], 32, 64);
        const R dr = Dv[2 * j], di = Dv[2 * j + 1];
        const R u = him ? fma(y[q], dr, yo * di) : fma(y[q], dr, -(yo * di));
        Lr[2 * (i * (PMAX + 1) + j) + (him ? 1 : 0)] = ok ? x[q] : R(0);
        Ur[2 * (i * (PMAX + 1) + j) + (him ? 1 : 0)] = ok ? u : R(0);
    }
    __syncthreads();
    inverse_product(Lis, Uis, t, NTHR, LUg);
}

// lane k of this lane's half-wave (lanes 0-31 or 32-63)
template <typename C> __device__ inline C shfl_half(C v, int k) {
    const int src = (int)(threadIdx.x & 32) + k;
    return mk<C>(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// ---------------------------------------------------------------------------
// Value of the second panel of a fused pair (first panel K1 = [k1, k1 + PMAX),
// second K2) at a row r that is no pivot of the first panel, column c of K2:
//   V2[r][c] = A[r][c] - sum_k V1[r][k] Z1[k][c],
// the V1 row (PMAX entries) and the tile Z1[:, K2] (zt[k][c], zero past the
// panel) in LDS.  The window panel and the check blocks of the Z build use
// this one routine, so both see the same values.
template <typename C>
__device__ inline C pair_second_value(C a, const C* v1row, const C (*zt)[PMAX + 1], int c) {
    C w = a;
#pragma unroll
    for (int k = 0; k < PMAX; k += 2) {
        const C va = v1row[k], vb = v1row[k + 1];
        w = csub(w, cmul(va, zt[k][c]));
        w = csub(w, cmul(vb, zt[k + 1][c]));
    }
    return w;
}

// ---------------------------------------------------------------------------
// Fast panel: GEPP on the candidate window, the first WIN unpivoted rows by
// index, by wave 0 (lane l owns window row win[l]); the whole workgroup
// (WP_THREADS) loads the window and forms the inverse factors.  Every chosen
// pivot is then checked against all rows outside the window (k_zbuild):
// threshold partial pivoting, |pivot| >= PIV_U * max |candidate| over every
// unpivoted row, or verr is raised and the factorization is redone with the
// full-column search.
// Outputs: pivots, V[piv, :]^-1 and U^-1 of the pivot block, updated window
// list.  The window's rows are read through LDS (lanes across columns,
// coalesced) before each lane of wave 0 takes its row into registers.
// ---------------------------------------------------------------------------
constexpr int WP_THREADS = 512;

template <typename R, int NB, bool SECOND>
__global__ __launch_bounds__(WP_THREADS, 2) void k_wpanel(const LineJob* jobs, int k0, int nb,
                                                        int fresh) {
    constexpr bool second = SECOND;  // second panel of a pair
    using C = cplx<R>;
    const LineJob J = jobs[blockIdx.x];
    const int n = J.n;
    const int t = threadIdx.x, lane = t & 63;
    const bool w0 = t < 64;
    const int nwin = min(WIN, n - k0);
    __shared__ int prs[NB];
    __shared__ R pmags[NB];
    __shared__ C Ls[PMAX][PMAX + 1];
    __shared__ C Us[PMAX][PMAX + 1];
    __shared__ C Ws[WIN][NB + 1];
    __shared__ int wrow[WIN];
    static_assert(NB == PMAX && WIN == 2 * PMAX, "window and panel shapes");
    static_assert(WIN * (NB + 1) >= 2 * PMAX * (PMAX + 1), "the inverse scratch reuses the window");
    if (nb < PMAX) {  // the inverses need L and U zero outside nb x nb
        for (int e = t; e < PMAX * (PMAX + 1); e += WP_THREADS) {
            (&Ls[0][0])[e] = mk<C>(R(0), R(0));
            (&Us[0][0])[e] = mk<C>(R(0), R(0));
        }
    }
    int row = -1;
    if (w0) {
        row = (lane < nwin) ? gload_s(J.win + lane) : -1;
        wrow[lane] = row;
    }
    bool act = row >= 0;
    int mystep = -1;  // panel step at which this lane's row became the pivot
    // panel values: the block's columns, or (second panel of a pair) the
    // columns as the first panel leaves them: the window's rows are no
    // pivots of the first panel, so V2 = A[:, K2] - V1 Z1[:, K2]
    const int poff = second ? PMAX : 0;
    __syncthreads();
    {
        // 32 lanes per row, WIN / 16 rows per thread; all loads in flight
        // before the first LDS store (lanes past the window or the panel
        // read entry 0 and store zero)
        constexpr int RPT = WIN * 32 / WP_THREADS;
        C v[RPT];
        const int c = t & 31, r0 = t >> 5;
        const C* Y = (const C*)J.Y;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int rw = wrow[r0 + 16 * q];
            v[q] = gload(Y + ((rw >= 0 && c < nb) ? (long)rw * n + k0 + c : 0));
        }
        if constexpr (second) {
            const C* Vb = (const C*)J.Vbuf;
            const C* Zg = (const C*)J.Z;
            // V1 rows of the window into the window tile, Z1[:, K2] into the
            // (not yet used) L factor tile
            constexpr int ZPT = PMAX * PMAX / WP_THREADS;
            C v1[RPT], zt[ZPT];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int rw = wrow[r0 + 16 * q];
                v1[q] = gload(Vb + (rw >= 0 ? (long)rw * VMAX + c : 0));
            }
#pragma unroll
            for (int q = 0; q < ZPT; ++q) {
                const int e = t + WP_THREADS * q, k = e / PMAX, cc = e % PMAX;
                zt[q] = gload(Zg + (cc < nb ? (long)k * n + k0 + cc : 0));
            }
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int rw = wrow[r0 + 16 * q];
                Ws[r0 + 16 * q][c] = csel(rw >= 0, v1[q], mk<C>(R(0), R(0)));
                v[q] = csel(c < nb, v[q], mk<C>(R(0), R(0)));
            }
#pragma unroll
            for (int q = 0; q < ZPT; ++q) {
                const int e = t + WP_THREADS * q, k = e / PMAX, cc = e % PMAX;
                Ls[k][cc] = csel(cc < nb, zt[q], mk<C>(R(0), R(0)));
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RPT; ++q) v[q] = pair_second_value(v[q], &Ws[r0 + 16 * q][0], Ls, c);
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int rw = wrow[r0 + 16 * q];
            Ws[r0 + 16 * q][c] = csel(rw >= 0 && c < nb, v[q], mk<C>(R(0), R(0)));
        }
    }
    __syncthreads();
    {
        // GEPP on the window by all waves: lane l of every wave holds window
        // row wrow[l], wave w its columns CPW w .. CPW w + CPW - 1, and the
        // steps run in blocks of CPW, block w by wave w.  The owner takes its
        // block's steps back to back on its own columns: at step cc it finds
        // the pivot (argmax of |.|_1 over the active rows; the lowest lane
        // among exact ties, i.e. the smallest row index, GEPP's tie rule),
        // forms the multipliers m = x / x[pivot] of the other active rows
        // (zero for the pivot and inactive rows), publishes them with the
        // pivot lane through LDS and updates its remaining columns.  After
        // one barrier per block (LDS double-buffered by block) every other
        // wave applies the block's steps to its own columns with the pivot
        // row's entries (readlane within the wave).  The arithmetic per entry
        // is that of a single-wave elimination, in the same order.  The wave
        // index is made wave-uniform so that the column tests are scalar
        // branches.
        constexpr int NW = WP_THREADS / 64, CPW = NB / NW;
        static_assert(CPW * NW == NB, "columns per wave");
        __shared__ C colb[2][CPW][WIN];
        __shared__ int plb[2][CPW];
        const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
        if (!w0) {
            row = wrow[lane];
            act = row >= 0;
        }
        C W[CPW];
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            const int c = CPW * wv + j;
            W[j] = mk<C>(R(0), R(0));
            if (act && c < nb) W[j] = Ws[lane][c];
        }
        const int nblk = (nb + CPW - 1) / CPW;
        for (int bk = 0; bk < nblk; ++bk) {
            const int bb = bk & 1;
            if (wv == bk) {
#pragma unroll
                for (int jj = 0; jj < CPW; ++jj) {
                    const int cc = CPW * bk + jj;
                    if (cc >= nb) break;  // uniform
                    const C x = W[jj];
                    const R mag = cabs1(x);
                    const auto key = act ? gepp_key(mag) : decltype(gepp_key(mag))(0);
                    const auto kmax = wave_max_key(key);
                    unsigned long long hit = __ballot(act && key == kmax);
                    if (hit == 0ULL) hit = __ballot(act);  // (cannot happen: some row is active)
                    const int pl = __builtin_amdgcn_readfirstlane((int)__ffsll((long long)hit) - 1);
                    const C ucc = mk<C>(bcast(x.x, pl), bcast(x.y, pl));
                    const bool live = act && lane != pl && cabs1(ucc) > R(0);
                    const C m = csel(live, cmul(x, crcp_scaled(ucc)), mk<C>(R(0), R(0)));
                    colb[bb][jj][lane] = m;
                    if (lane == 0) plb[bb][jj] = pl;
                    if (lane == pl) {
                        prs[cc] = row;
                        pmags[cc] = act ? mag : R(-1);
                        act = false;
                        mystep = cc;
                    }
                    // the pivot lane's registers are not modified in this step
                    W[jj] = csel(act, m, W[jj]);
#pragma unroll
                    for (int j = jj + 1; j < CPW; ++j) {
                        const C uc = mk<C>(bcast(W[j].x, pl), bcast(W[j].y, pl));
                        W[j] = csel(act, csub(W[j], cmul(m, uc)), W[j]);
                    }
                }
            }
            __syncthreads();
            if (wv != bk) {
                const int nst = min(CPW, nb - CPW * bk);
                int pls[CPW];
                C ms[CPW];
#pragma unroll
                for (int jj = 0; jj < CPW; ++jj) {
                    pls[jj] = plb[bb][jj];
                    ms[jj] = colb[bb][jj][lane];
                }
#pragma unroll
                for (int jj = 0; jj < CPW; ++jj) {
                    if (jj >= nst) break;  // uniform
                    const int pl = __builtin_amdgcn_readfirstlane(pls[jj]);
                    if (lane == pl) {
                        act = false;
                        mystep = CPW * bk + jj;
                    }
                    // columns right of the block are brought up to date;
                    // those left of it hold final multipliers
                    if (wv > bk) {
#pragma unroll
                        for (int j = 0; j < CPW; ++j) {
                            const C uc = mk<C>(bcast(W[j].x, pl), bcast(W[j].y, pl));
                            W[j] = csel(act, csub(W[j], cmul(ms[jj], uc)), W[j]);
                        }
                    }
                }
            }
        }
        // A pivot lane's registers are frozen from its step on: they hold its
        // row of L (multipliers, c < step) and of U (c >= step).
        if (mystep >= 0) {
#pragma unroll
            for (int j = 0; j < CPW; ++j) {
                const int c = CPW * wv + j;
                if (c < nb) {
                    const bool lo = c < mystep;
                    Ls[mystep][c] = csel(lo, W[j], mk<C>(R(0), R(0)));  // strictly lower
                    Us[mystep][c] = csel(lo, mk<C>(R(0), R(0)), W[j]);
                }
            }
        }
    }
    __syncthreads();
    // the window tile is dead: its LDS holds the inverse scratch
    C(*Lis)[PMAX + 1] = (C(*)[PMAX + 1]) & Ws[0][0];
    __shared__ R dvs[2 * PMAX];  // 1 / U[i][i]
    store_pivot_inverses_w<WP_THREADS, C>(Ls, Us, Lis, Lis + PMAX, dvs, nb, t, (C*)J.LU);
    if (!w0) return;
    if (lane < nb) {
        const int p = prs[lane];
        J.piv[poff + lane] = p;
        J.prow[k0 + lane] = p;
        J.iprow[p] = k0 + lane;
        J.used[p] = 1;
    }
    if (lane == 0) {
        double mn = J.pivstat[0], mx = J.pivstat[1];
        bool zero = false;
        for (int c = 0; c < nb; ++c) {
            double v = (double)pmags[c];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
            if (!(v > 0)) zero = true;
        }
        J.pivstat[0] = mn;
        J.pivstat[1] = mx;
        // a zero (or NaN) pivot in the window is not final: the full-column
        // search redoes the factorization and decides
        if (zero) {
            J.pivstat[2] = 1.0;
            *(GLOBAL_AS int*)J.verr = 1;
        }
    }
    // new window: surviving rows, then the next nb fresh rows
    const bool keep = (row >= 0) && mystep < 0;
    const unsigned long long km = __ballot(keep);
    const int below = __popcll(km & ((1ull << lane) - 1ull));
    const int nkeep = __popcll(km);
    if (keep) J.win[below] = row;
    if (lane < nb && fresh + lane < n) J.win[nkeep + lane] = fresh + lane;
}

template <typename R>
void launch_wpanel(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb,
                   int fresh, int second) {
    (void)n;
    if (second) k_wpanel<R, PMAX, true><<<njobs, WP_THREADS, 0, s>>>(jobs_d, k0, nb, fresh);
    else k_wpanel<R, PMAX, false><<<njobs, WP_THREADS, 0, s>>>(jobs_d, k0, nb, fresh);
}
template void launch_wpanel<double>(hipStream_t, const LineJob*, int, int, int, int, int, int);
template void launch_wpanel<float>(hipStream_t, const LineJob*, int, int, int, int, int, int);

// ---------------------------------------------------------------------------
// Full panel (fallback, exact GEPP over every unpivoted row): one workgroup of
// 1024 threads, thread t owns row t.  NB is a compile-time width (16 .. 1) so
// the row segment stays in registers.  Outputs as k_wpanel plus the panel copy.
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
    __shared__ int prs[NB];
    __shared__ R pmags[NB];
    __shared__ C Ls[PMAX][PMAX + 1];
    __shared__ C Us[PMAX][PMAX + 1];
    __shared__ C Lis[2 * PMAX][PMAX + 1];  // inverse scratch (L^-1, U^-1)
    const C* A = (const C*)J.Y;
    const int r = t;
    const bool valid = r < n;
    bool act = valid && (J.used[r] == 0);
    const C zero = mk<C>(R(0), R(0));
    C W[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) W[c] = valid ? A[(long)r * n + k0 + c] : mk<C>(R(0), R(0));
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
            const C v = cval[buf][pw][t];
            Ls[cc][t] = csel(t < cc, v, mk<C>(t == cc ? R(1) : R(0), R(0)));
            Us[cc][t] = csel(t < cc, zero, v);
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
    store_pivot_inverses<C>(Ls, Us, Lis, Lis + PMAX, NB, t, 1024, (C*)J.LU);
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
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
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
// Z build, panel copy and pivot check, on MFMA (16 x 16 x 4 products, complex
// values as four real products).  M = V[piv, :]^-1 and U^-1 come from the
// panel kernel.
//   blocks [0, nzb):  64 columns per workgroup, 16 per wave:
//                     T[k][col] = A[piv[k]][col] (delta(k, col - k0) for col
//                     in K), Z[:, col] = M T[:, col];
//   blocks [nzb, ..): 32 rows per workgroup: copy V = A[r, K] to the panel
//                     buffer and, if verify, check rows r >= fresh: the value
//                     GEPP would compare at step c is a_c = y_c U[c][c] with
//                     y = V[r] U^-1, and PIV_U |a_c| <= |U[c][c]| must hold
//                     (threshold pivoting).
// Second panel of a fused pair (columns K2 after K1 = [k1, k1 + nb1)): its
// pivot rows are taken after the first panel's update,
//   T[k][col] = [col not in K1] A[p2_k][col] - sum_q V1[p2_k][q] Z1[q][col],
// Z goes to rows PMAX.. of the Z buffer, and the check blocks form the
// panel's values for every row as the first panel leaves them (the first
// panel's pivots take their Z1 row, every other row A - V1 Z1[:, K2]), write
// them to columns PMAX.. of the panel buffer and check them.
// Operand layout of v_mfma_*_16x16x4: A[i = lane%16][k = lane/16],
// B[k = lane/16][j = lane%16], result D[M::row(lane, v)][lane%16].  In Z = M T
// the contraction index of step (rb, s) for lane group g is the row T holds
// in its accumulator register s, rb 16 + M::row(16 g, s), so T feeds the
// product straight from the registers it was accumulated in.
// ---------------------------------------------------------------------------
constexpr int ZB_COLS = 64;  // columns per Z workgroup

// SECOND: the second panel of a pair (a template parameter, so that the
// first panel's launches carry none of its registers)
template <typename R, bool SECOND>
__global__ __launch_bounds__(256) void k_zbuild(const LineJob* jobs, int k0, int nb, int fresh,
                                                int verify, int nzb, int k1, int nb1) {
    constexpr bool second = SECOND;
    using C = cplx<R>;
    using M = mfma_t<R>;
    using acc_t = typename M::acc;
    const LineJob J = jobs[blockIdx.y];
    const int n = J.n;
    // M (Z blocks) or U^-1 (check blocks); in the check blocks of a second
    // panel first the tile Z1[:, K2]
    __shared__ __align__(16) R Fri[2][PMAX][PMAX + 1];
    R(*Fr)[PMAX + 1] = Fri[0];
    R(*Fi)[PMAX + 1] = Fri[1];
    // V1[piv2, :] (Z) or the V tile (check); in the check blocks of a second
    // panel first the rows' V1 as one complex tile
    __shared__ __align__(16) R Vri[2][PMAX][PMAX + 1];
    R(*Vr)[PMAX + 1] = Vri[0];
    R(*Vi)[PMAX + 1] = Vri[1];
    __shared__ int pv[PMAX];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const bool zpart = (int)blockIdx.x < nzb;
    const int poff = second ? PMAX : 0;
    constexpr int EPT = PMAX * PMAX / 256;  // factor entries per thread (blockDim.x == 256)
    // check blocks: the V tile (rows rbase.., panel columns), 32 lanes per
    // row, issued before the factor staging so the two latencies overlap;
    // out-of-range lanes read the first panel entry and are zeroed
    const int vc = t & 31, h8 = t >> 5;
    const int rbase = ((int)blockIdx.x - nzb) * 32;
    C vrow[4];
    // second panel of a pair: the check blocks form the panel's values as
    // the first panel leaves them (its pivots take their Z1 row, every other
    // row A - V1 Z1[:, K2]) and write them to the panel buffer
    C v1r[4];
    int slot[4];
    if (!zpart) {
        const C* Y = (const C*)J.Y;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = rbase + h8 + 8 * i;
            const bool ok = r < n && vc < nb;
            vrow[i] = gload(Y + (ok ? (long)r * n + k0 + vc : 0));
        }
        if constexpr (second) {
            const C* Vb = (const C*)J.Vbuf;
            const C* Zg = (const C*)J.Z;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = rbase + h8 + 8 * i;
                slot[i] = gload_s(J.iprow + (r < n ? r : 0)) - k1;
                v1r[i] = gload(Vb + (r < n ? (long)r * VMAX + vc : 0));
            }
        }
    }
    // check blocks of a second panel: U^-1 waits in registers while the
    // factor tile holds Z1[:, K2]
    const bool pair2 = second && !zpart;
    C fv[EPT];
    {
        // M or U^-1 and the pivot rows: every load issued before the first use
        const C* Fg = (const C*)J.LU + (zpart ? 0 : PMAX * PMAX);
        const int p = gload_s(J.piv + poff + (t < nb ? t : 0));
#pragma unroll
        for (int q = 0; q < EPT; ++q) fv[q] = gload(Fg + t + 256 * q);
        if (t < PMAX) pv[t] = (t < nb) ? p : 0;  // rows past nb: a safe row, masked
        if (pair2) {
            C zt[EPT];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = t + 256 * q, k = e / PMAX, cc = e % PMAX;
                zt[q] = gload((const C*)J.Z + (cc < nb ? (long)k * n + k0 + cc : 0));
            }
            C(*Z1s)[PMAX + 1] = (C(*)[PMAX + 1]) & Fri[0][0][0];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = t + 256 * q, k = e / PMAX, cc = e % PMAX;
                Z1s[k][cc] = csel(cc < nb, zt[q], mk<C>(R(0), R(0)));
            }
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = t + 256 * q;
                Fr[e / PMAX][e % PMAX] = fv[q].x;
                Fi[e / PMAX][e % PMAX] = fv[q].y;
            }
        }
    }
    __syncthreads();
    const C* A = (const C*)J.Y;
    C* Zg = (C*)J.Z;
    const C* Vb = (const C*)J.Vbuf;
    if (zpart) {
        if constexpr (second) {
            C v[EPT];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = t + 256 * q, k = e / PMAX, c = e % PMAX;
                v[q] = gload(Vb + ((k < nb && c < nb1) ? (long)pv[k] * VMAX + c : 0));
            }
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = t + 256 * q, k = e / PMAX, c = e % PMAX;
                const bool ok = k < nb && c < nb1;
                Vr[k][c] = ok ? v[q].x : R(0);
                Vi[k][c] = ok ? v[q].y : R(0);
            }
            __syncthreads();
        }
        const int col = blockIdx.x * ZB_COLS + wv * 16 + lr;
        const bool colok = col < n;
        const bool inK = col >= k0 && col < k0 + nb;
        const bool inK1 = second && col >= k1 && col < k1 + nb1;
        const bool ldA = colok && !inK && !inK1;
        // T in accumulator layout: register v of row block rb holds
        // T[rb 16 + M::row(lane, v)][col]
        acc_t tre[2], tim[2];
        {
            C tv[2][4];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = rb * 16 + M::row(lane, v);
                    tv[rb][v] = gload(A + ((ldA && k < nb) ? (long)pv[k] * n + col : 0));
                }
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = rb * 16 + M::row(lane, v);
                    const bool id = inK && k < nb && col - k0 == k;
                    tre[rb][v] = (ldA && k < nb) ? tv[rb][v].x : (id ? R(1) : R(0));
                    tim[rb][v] = (ldA && k < nb) ? tv[rb][v].y : R(0);
                }
        }
        if constexpr (second) {
            // T -= V1[piv2, :] Z1 (columns outside K2): A = V1[k][q], B = Z1[q][col],
            // q = 4 s + lane group
            const bool zok = colok && !inK;
            C zq[PMAX / 4];
#pragma unroll
            for (int st = 0; st < PMAX / 4; ++st) {
                const int q = 4 * st + lg;
                zq[st] = gload(Zg + ((zok && q < nb1) ? (long)q * n + col : 0));
            }
#pragma unroll
            for (int st = 0; st < PMAX / 4; ++st) {
                const int q = 4 * st + lg;
                const bool ok = zok && q < nb1;
                const R br = ok ? zq[st].x : R(0), bi = ok ? zq[st].y : R(0);
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    const R ar = Vr[rb * 16 + lr][q], ai = Vi[rb * 16 + lr][q];
                    tre[rb] = M::run(-ar, br, tre[rb]);
                    tre[rb] = M::run(ai, bi, tre[rb]);
                    tim[rb] = M::run(-ar, bi, tim[rb]);
                    tim[rb] = M::run(-ai, br, tim[rb]);
                }
            }
        }
        // Z = M T
        acc_t zre[2], zim[2];
#pragma unroll
        for (int ro = 0; ro < 2; ++ro) {
            zre[ro] = acc_t{0, 0, 0, 0};
            zim[ro] = acc_t{0, 0, 0, 0};
        }
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const int k = rb * 16 + M::row(lane, st);
                const R br = tre[rb][st], bi = tim[rb][st];
#pragma unroll
                for (int ro = 0; ro < 2; ++ro) {
                    const R ar = Fr[ro * 16 + lr][k], ai = Fi[ro * 16 + lr][k];
                    zre[ro] = M::run(ar, br, zre[ro]);
                    zre[ro] = M::run(-ai, bi, zre[ro]);
                    zim[ro] = M::run(ar, bi, zim[ro]);
                    zim[ro] = M::run(ai, br, zim[ro]);
                }
            }
        if (!colok) return;
#pragma unroll
        for (int ro = 0; ro < 2; ++ro)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = ro * 16 + M::row(lane, v);
                if (row < nb) gstore(Zg + (long)(poff + row) * n + col, mk<C>(zre[ro][v], zim[ro][v]));
            }
    } else {
        const int c = vc;
        if constexpr (second) {
            C(*V1s)[PMAX + 1] = (C(*)[PMAX + 1]) & Vri[0][0][0];
#pragma unroll
            for (int i = 0; i < 4; ++i) V1s[h8 + 8 * i][c] = (c < nb1) ? v1r[i] : mk<C>(R(0), R(0));
            const C(*Z1s)[PMAX + 1] = (const C(*)[PMAX + 1]) & Fri[0][0][0];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // the first panel's pivots take their row of the Z1 tile
                const int sl = slot[i];
                const bool piv1 = sl >= 0 && sl < nb1;
                const C a = csel(c < nb, vrow[i], mk<C>(R(0), R(0)));
                vrow[i] = piv1 ? Z1s[piv1 ? sl : 0][c] : pair_second_value(a, &V1s[h8 + 8 * i][0], Z1s, c);
            }
            __syncthreads();  // both tiles are read: their LDS takes U^-1 and the V tile
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                const int e = t + 256 * q;
                Fr[e / PMAX][e % PMAX] = fv[q].x;
                Fi[e / PMAX][e % PMAX] = fv[q].y;
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = rbase + h8 + 8 * i;
            const bool ok = r < n && c < nb;
            const C v = csel(ok, vrow[i], mk<C>(R(0), R(0)));
            if (ok) gstore((C*)J.Vbuf + (long)r * VMAX + poff + c, v);
            Vr[h8 + 8 * i][c] = v.x;
            Vi[h8 + 8 * i][c] = v.y;
        }
        if (!verify) return;  // uniform
        __syncthreads();
        // Y = V U^-1, wave (rb, cb) -> rows rb 16.., columns cb 16..
        const int rb = wv >> 1, cb = wv & 1;
        acc_t yre = acc_t{0, 0, 0, 0}, yim = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int st = 0; st < PMAX / 4; ++st) {
            const int q = 4 * st + lg;
            const R ar = Vr[rb * 16 + lr][q], ai = Vi[rb * 16 + lr][q];
            const R br = Fr[q][cb * 16 + lr], bi = Fi[q][cb * 16 + lr];
            yre = M::run(ar, br, yre);
            yre = M::run(-ai, bi, yre);
            yim = M::run(ar, bi, yim);
            yim = M::run(ai, br, yim);
        }
        const int cc = cb * 16 + lr;
        bool bad = false;
        if (cc < nb) {
            const C ucc = crcp(mk<C>(Fr[cc][cc], Fi[cc][cc]));  // U[c][c] = 1 / U^-1[c][c]
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rbase + rb * 16 + M::row(lane, v);
                const C ac = cmul(mk<C>(yre[v], yim[v]), ucc);
                // (written so that a NaN, from a singular pivot block, also fails)
                if (r < n && r >= fresh && !(R(PIV_U) * cabs1(ac) <= cabs1(ucc))) bad = true;
            }
        }
        if (bad) *(GLOBAL_AS int*)J.verr = 1;
    }
}

template <typename R>
void launch_zbuild(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb,
                   int fresh, int verify, int second, int k1, int nb1) {
    const int nzb = (n + ZB_COLS - 1) / ZB_COLS;
    const int nvb = (n + 31) / 32;
    if (second)
        k_zbuild<R, true><<<dim3(nzb + nvb, njobs), 256, 0, s>>>(jobs_d, k0, nb, fresh, verify, nzb, k1, nb1);
    else
        k_zbuild<R, false><<<dim3(nzb + nvb, njobs), 256, 0, s>>>(jobs_d, k0, nb, fresh, verify, nzb, k1, nb1);
}
template void launch_zbuild<double>(hipStream_t, const LineJob*, int, int, int, int, int, int,
                                    int, int, int);
template void launch_zbuild<float>(hipStream_t, const LineJob*, int, int, int, int, int, int,
                                   int, int, int);

// ---------------------------------------------------------------------------
// Trailing update on MFMA.  64x64 output tile per block, 4 waves of 32x32,
// 16x16x4 MFMA blocks, complex product as four real MFMAs:
//   re += Vr Zr - Vi Zi,  im += Vr Zi + Vi Zr.
// Operand layout of v_mfma_*_16x16x4: A[i = lane%16][k = lane/16],
// B[k = lane/16][j = lane%16]; result D[mfma_t::row(lane, v)][lane%16].
// The old tile values are loaded before the MFMA loop so their latency
// overlaps the operand staging.
// ---------------------------------------------------------------------------

// Trailing update of one panel (nb2 = 0, columns K = [k0, k0 + nb1)):
//   A[i, :] = A'[i, :] - V[i, :] Z  (A' = A with K zeroed),  A[piv, :] = Z;
// or of a fused pair (K1 = [k0, k0 + PMAX), K2 = the next nb2 columns):
//   rows of neither panel:  A'' - V1 Z1' - V2 Z2   (A'' zeroes K1 and K2,
//                                                  Z1' = Z1 with K2 zeroed)
//   first-panel pivots:     Z1'[slot] - V2 Z2
//   second-panel pivots:    Z2[slot]
// The rank-(nb1 + nb2) product runs in chunks of PMAX columns of V through
// the same LDS tiles.
template <typename R>
__global__ __launch_bounds__(256, 3) void k_trailing(const LineJob* jobs, int k0, int nb1, int nb2) {
    using C = cplx<R>;
    using M = mfma_t<R>;
    using acc_t = typename M::acc;
    constexpr int NB = PMAX;
    const LineJob J = jobs[blockIdx.z];
    const int n = J.n;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;
    const bool fused = nb2 > 0;
    const int kend = fused ? k0 + PMAX + nb2 : k0 + nb1;  // end of the updated columns
    const int k2 = k0 + PMAX;                               // start of K2 (fused)
    // V and Z pass through LDS KH columns at a time (a quarter of the LDS of a
    // whole 32-column chunk, so more workgroups share a CU)
    constexpr int KH = 16;
    __shared__ R Vr[64][KH + 1], Vi[64][KH + 1];
    __shared__ R Zr[KH][64 + 1], Zi[KH][64 + 1];
    __shared__ int rslot[64];
    C* A = (C*)J.Y;
    const C* Vb = (const C*)J.Vbuf;
    const C* Zg = (const C*)J.Z;
    if (t < 64) rslot[t] = (i0 + t < n) ? gload_s(J.iprow + i0 + t) - k0 : -1;
    acc_t are[2][2], aim[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            are[bi][bj] = acc_t{0, 0, 0, 0};
            aim[bi][bj] = acc_t{0, 0, 0, 0};
        }
    __syncthreads();  // rslot
    const int nchunk = fused ? 2 : 1;
    bool first = true;
    for (int ch = 0; ch < nchunk; ++ch) {
        const int koff = ch * PMAX, nbc = (ch == 0) ? nb1 : nb2;
        for (int h0 = 0; h0 < nbc; h0 += KH) {
            if (!first) __syncthreads();  // the previous columns' tiles are consumed
            first = false;
            for (int e = t; e < 64 * KH; e += 256) {
                const int ii = e / KH, k = e % KH, i = i0 + ii;
                C v = mk<C>(R(0), R(0));
                // the first panel's pivot rows take no V1 part in a fused pair
                const bool skip = fused && ch == 0 && rslot[ii] >= 0 && rslot[ii] < PMAX;
                if (i < n && h0 + k < nbc && !skip) v = gload(Vb + (long)i * VMAX + koff + h0 + k);
                Vr[ii][k] = v.x;
                Vi[ii][k] = v.y;
            }
            for (int e = t; e < 64 * KH; e += 256) {
                const int k = e / 64, jj = e % 64, j = j0 + jj;
                C v = mk<C>(R(0), R(0));
                // Z1' : Z1 with the K2 columns zeroed
                const bool skip = fused && ch == 0 && j >= k2 && j < kend;
                if (j < n && h0 + k < nbc && !skip) v = gload(Zg + (long)(koff + h0 + k) * n + j);
                Zr[k][jj] = v.x;
                Zi[k][jj] = v.y;
            }
            __syncthreads();
            const int nbr = (min(KH, nbc - h0) + 3) & ~3;
            for (int kk = 0; kk < nbr; kk += 4) {
                R ar[2], an[2], ai[2], br[2], bi_[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    ar[q] = Vr[wr * 32 + q * 16 + lr][kk + lk];
                    ai[q] = Vi[wr * 32 + q * 16 + lr][kk + lk];
                    an[q] = -ai[q];
                    br[q] = Zr[kk + lk][wc * 32 + q * 16 + lr];
                    bi_[q] = Zi[kk + lk][wc * 32 + q * 16 + lr];
                }
                // passes over the four blocks, so consecutive MFMAs use
                // different accumulators (each accumulator's order unchanged)
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) are[bi][bj] = M::run(ar[bi], br[bj], are[bi][bj]);
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) aim[bi][bj] = M::run(ar[bi], bi_[bj], aim[bi][bj]);
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) are[bi][bj] = M::run(an[bi], bi_[bj], are[bi][bj]);
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) aim[bi][bj] = M::run(ai[bi], br[bj], aim[bi][bj]);
            }
        }
    }
    // old tile values and the pivot-row slots (loaded after the products, so
    // they hold no registers during them; other workgroups cover the latency)
    C old[2][4][2];
    int slot[2][4];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + wr * 32 + bi * 16 + M::row(lane, v);
            const int sl = (i < n) ? gload_s(J.iprow + i) - k0 : -1;
            slot[bi][v] = sl;
            const bool p1 = fused && sl >= 0 && sl < PMAX;
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int j = j0 + wc * 32 + bj * 16 + lr;
                old[bi][v][bj].x = R(0);
                old[bi][v][bj].y = R(0);
                if (i < n && j < n) {
                    if (p1) {
                        if (!(j >= k2 && j < kend)) old[bi][v][bj] = gload(Zg + (long)sl * n + j);
                    } else if (!(j >= k0 && j < kend)) {
                        old[bi][v][bj] = gload(A + (long)i * n + j);
                    }
                }
            }
        }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + wr * 32 + bi * 16 + M::row(lane, v);
            if (i >= n) continue;
            const int sl = slot[bi][v];
            // rows replaced by Z: the panel's pivots (single) or the second
            // panel's pivots (fused)
            const bool zrow = fused ? (sl >= PMAX && sl < PMAX + nb2) : (sl >= 0 && sl < nb1);
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int jj = wc * 32 + bj * 16 + lr;
                const int j = j0 + jj;
                if (j >= n) continue;
                C out;
                if (zrow) out = gload(Zg + (long)sl * n + j);
                else out = mk<C>(old[bi][v][bj].x - are[bi][bj][v], old[bi][v][bj].y - aim[bi][bj][v]);
                gstore(A + (long)i * n + j, out);
            }
        }
}

// FP32 trailing update for lines with n % 4 == 0: the same product as
// k_trailing<float> with the operands of the MFMA swapped (A = Z^T, B = V^T),
// so each lane holds one row i and four consecutive columns j = 4 (lane / 16)
// + v of the result and reads and writes them as two 16-byte accesses.
//   * V and Z tiles are fetched as 16-byte pairs into registers without
//     branches (out-of-range lanes read offset 0; a bit mask applied when
//     the pair is stored to LDS zeroes them), so a fused pair's second chunk
//     is in flight while the first chunk's MFMAs run.
//   * The old tile values are fetched after chunk 0 is staged; the row slots
//     come from LDS, so no global load forces an early wait.
//   * The flat grid is ordered so consecutive tiles of a line land on one
//     XCD (workgroups are dealt to the 8 XCDs round-robin) and share its L2.
// n % 4 == 0 keeps every 16-byte access aligned and inside its row.
__global__ __launch_bounds__(256) void k_trailing_f32(const LineJob* jobs, int k0, int nb1, int nb2,
                                                      int ntx, int total) {
    using R = float;
    using C = cplx<R>;
    using M = mfma_t<R>;
    using acc_t = typename M::acc;
    constexpr int NB = PMAX;
    // XCD-aware tile order: the first 8 * per workgroups are dealt so that
    // XCD x (= L % 8) runs the contiguous tile range [x per, (x + 1) per)
    const int L = blockIdx.x, per = total / 8;
    const int lin = (L < 8 * per) ? (L % 8) * per + L / 8 : L;
    const int job = lin / (ntx * ntx), rem = lin % (ntx * ntx);
    const LineJob J = jobs[job];
    const int n = J.n;
    const int i0 = (rem / ntx) * 64, j0 = (rem % ntx) * 64;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;
    const bool fused = nb2 > 0;
    const int kend = fused ? k0 + PMAX + nb2 : k0 + nb1;  // end of the updated columns
    const int k2 = k0 + PMAX;                               // start of K2 (fused)
    __shared__ R Vr[64][NB + 1], Vi[64][NB + 1];
    __shared__ R Zr[NB][64 + 1], Zi[NB][64 + 1];
    __shared__ int rslot[64];
    C* A = (C*)J.Y;
    const C* Vb = (const C*)J.Vbuf;
    const C* Zg = (const C*)J.Z;
    if (t < 64) rslot[t] = (i0 + t < n) ? gload_s(J.iprow + i0 + t) - k0 : -1;
    acc_t are[2][2], aim[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            are[bi][bj] = acc_t{0, 0, 0, 0};
            aim[bi][bj] = acc_t{0, 0, 0, 0};
        }
    __syncthreads();  // rslot
    // element pairs: V[ii][kp, kp + 1] with ii = e / 16, kp = 2 (e % 16);
    // Z[k][jp, jp + 1] with k = e / 32, jp = 2 (e % 32); e = t + 256 q
    f4v pv[4], pz[4];
    unsigned msk = 0;  // bits 2q, 2q+1: V pair q halves; bits 8+2q, 9+2q: Z pair q halves
    auto fetch = [&](int ch) {
        const int koff = ch * PMAX, nbc = (ch == 0) ? nb1 : nb2;
        msk = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int ii = e >> 4, kp = (e & 15) * 2;
            const int i = i0 + ii;
            // the first panel's pivot rows take no V1 part in a fused pair
            const bool skip = fused && ch == 0 && rslot[ii] >= 0 && rslot[ii] < PMAX;
            const bool okv = i < n && !skip;
            msk |= (unsigned)(okv && kp < nbc) << (2 * q);
            msk |= (unsigned)(okv && kp + 1 < nbc) << (2 * q + 1);
            pv[q] = gload2c(Vb + (okv ? (long)i * VMAX + koff + kp : 0));
            const int k = e >> 5, j = j0 + (e & 31) * 2;
            const bool okz = k < nbc && j < n;
            // Z1' : Z1 with the K2 columns zeroed
            const bool ex0 = fused && ch == 0 && j >= k2 && j < kend;
            const bool ex1 = fused && ch == 0 && j + 1 >= k2 && j + 1 < kend;
            msk |= (unsigned)(okz && !ex0) << (8 + 2 * q);
            msk |= (unsigned)(okz && !ex1) << (9 + 2 * q);
            pz[q] = gload2c(Zg + (okz ? (long)(koff + k) * n + j : 0));
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int ii = e >> 4, kp = (e & 15) * 2;
            const bool a0 = (msk >> (2 * q)) & 1, a1 = (msk >> (2 * q + 1)) & 1;
            Vr[ii][kp] = a0 ? pv[q][0] : R(0);
            Vi[ii][kp] = a0 ? pv[q][1] : R(0);
            Vr[ii][kp + 1] = a1 ? pv[q][2] : R(0);
            Vi[ii][kp + 1] = a1 ? pv[q][3] : R(0);
            const int k = e >> 5, jp = (e & 31) * 2;
            const bool b0 = (msk >> (8 + 2 * q)) & 1, b1 = (msk >> (9 + 2 * q)) & 1;
            Zr[k][jp] = b0 ? pz[q][0] : R(0);
            Zi[k][jp] = b0 ? pz[q][1] : R(0);
            Zr[k][jp + 1] = b1 ? pz[q][2] : R(0);
            Zi[k][jp + 1] = b1 ? pz[q][3] : R(0);
        }
    };
    // old values of this lane's row i, columns jb .. jb + 3 (two 16-byte
    // pairs); exclusions are applied in the epilogue
    f4v raw[2][2][2];
    int slot[2];
    auto load_old = [&]() {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int ii = wr * 32 + bi * 16 + lr, i = i0 + ii;
            const int sl = rslot[ii];
            slot[bi] = sl;
            const bool p1 = fused && sl >= 0 && sl < PMAX;
            const C* src = p1 ? Zg : A;
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int jb = j0 + wc * 32 + bj * 16 + 4 * lk;
                const long off = (i < n && jb < n) ? (p1 ? (long)sl * n + jb : (long)i * n + jb) : 0;
                raw[bi][bj][0] = gload2c(src + off);
                raw[bi][bj][1] = gload2c(src + off + 2);
            }
        }
    };
    fetch(0);
    const int nchunk = fused ? 2 : 1;
    for (int ch = 0; ch < nchunk; ++ch) {
        const int nbc = (ch == 0) ? nb1 : nb2;
        if (ch > 0) __syncthreads();  // the previous chunk's tiles are consumed
        stash();
        if (ch == 0) {
            if (fused) fetch(1);
            load_old();
        }
        __syncthreads();
        const int nbr = (nbc + 3) & ~3;
        for (int kk = 0; kk < nbr; kk += 4) {
            R ar[2], ai[2], br[2], bi_[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                ar[q] = Vr[wr * 32 + q * 16 + lr][kk + lk];
                ai[q] = Vi[wr * 32 + q * 16 + lr][kk + lk];
                br[q] = Zr[kk + lk][wc * 32 + q * 16 + lr];
                bi_[q] = Zi[kk + lk][wc * 32 + q * 16 + lr];
            }
            // D[j][i] += sum_k Z[k][j] V[i][k]: result row = column j of the tile
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    are[bi][bj] = M::run(br[bj], ar[bi], are[bi][bj]);
                    are[bi][bj] = M::run(-bi_[bj], ai[bi], are[bi][bj]);
                    aim[bi][bj] = M::run(bi_[bj], ar[bi], aim[bi][bj]);
                    aim[bi][bj] = M::run(br[bj], ai[bi], aim[bi][bj]);
                }
        }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int i = i0 + wr * 32 + bi * 16 + lr;
        if (i >= n) continue;
        const int sl = slot[bi];
        const bool p1 = fused && sl >= 0 && sl < PMAX;
        // rows replaced by Z: the panel's pivots (single) or the second
        // panel's pivots (fused)
        const bool zrow = fused ? (sl >= PMAX && sl < PMAX + nb2) : (sl >= 0 && sl < nb1);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int jb = j0 + wc * 32 + bj * 16 + 4 * lk;
            if (jb >= n) continue;
            R o[8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = jb + v;
                const bool ex = p1 ? (j >= k2 && j < kend) : (j >= k0 && j < kend);
                const f4v& rw = raw[bi][bj][v >> 1];
                const R orr = ex ? R(0) : rw[(v & 1) * 2], oii = ex ? R(0) : rw[(v & 1) * 2 + 1];
                o[2 * v] = orr - are[bi][bj][v];
                o[2 * v + 1] = oii - aim[bi][bj][v];
            }
            f4v o0 = {o[0], o[1], o[2], o[3]}, o1 = {o[4], o[5], o[6], o[7]};
            if (zrow) {
                o0 = gload2c(Zg + (long)sl * n + jb);
                o1 = gload2c(Zg + (long)sl * n + jb + 2);
            }
            gstore2c(A + (long)i * n + jb, o0);
            gstore2c(A + (long)i * n + jb + 2, o1);
        }
    }
}

// Mixed-path trailing update on fp16x3 products (n % 4 == 0): the update,
// exclusions and pivot-row rules of k_trailing_f32 with the products formed
// in fp16x3.  Every chunk of at most 32 pivot columns is one K step of
// v_mfma_f32_16x16x32_f16, so the workgroup splits the operands itself: the
// V rows and Z columns are scaled by 2^split_exp(largest |re|, |im| over both
// chunks) and written to LDS as fp16 hi / lo planes in the fragment layout
// (lane l of a 16-row block reads the 16 bytes at 16 l: row l % 16,
// k = 8 (l / 16) + [0, 8)); each real product is hi hi + hi lo + lo hi with
// FP32 accumulation, unscaled exactly by 2^-(e_i + e_j).  The products go
// through LDS so that the old tile is read and the result written as whole
// 512-byte rows (row-contiguous tiles stream at about 1.5x the rate of the
// MFMA output layout's 32-byte pieces); a row's source (A, or its Z row for
// pivot rows) is chosen when it is loaded, so the epilogue issues no
// dependent load.
//
// One workgroup per 64-row strip and TPW consecutive 64 x 64 tiles of it:
// the strip's V rows are split once (with the first tile's loads already in
// flight), and while a tile's products and epilogue run the next tile's Z
// columns are already loading; its old rows are issued right after the
// epilogue.  LDS: the V planes (32 KB, kept for
// the strip) and one 36 KB region used in turn by a tile's Z planes and its
// product tile (two workgroups per CU).
constexpr int H3_PROW = 66;  // product tile row stride (complex): 528 bytes
constexpr int H3_TPW = 4;    // tiles per workgroup (2, 8 and 16 measured slower)

template <int TPW>
__global__ __launch_bounds__(256, 2) void k_trailing_h3s(const LineJob* jobs, int k0, int nb1,
                                                         int nb2, int nts, int ngrp, int total) {
    using C = float2;
    const int L = blockIdx.x, per = total / 8;
    const int lin = (L < 8 * per) ? (L % 8) * per + L / 8 : L;
    const int job = lin / (nts * ngrp), rem = lin % (nts * ngrp);
    const LineJob J = jobs[job];
    const int n = J.n;
    const int i0 = (rem / ngrp) * 64, tj0 = (rem % ngrp) * TPW;
    const int ntl = min(TPW, nts - tj0);  // tiles of this workgroup
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int lr = lane & 15, lk = lane >> 4;
    const bool fused = nb2 > 0;
    const int nchunk = fused ? 2 : 1;
    const int kend = fused ? k0 + PMAX + nb2 : k0 + nb1;  // end of the updated columns
    const int k2 = k0 + PMAX;                               // start of K2 (fused)
    __shared__ __align__(16) h8v sV[2][4][4][64];  // [chunk][plane][16-row block][lane]
    __shared__ __align__(16) h8v sZP[2][4][4][64 + 8];  // Z planes, then the product tile
    __shared__ int eV[2][64], eZ[2][64];  // per chunk
    h8v (*sZ)[4][4][72] = sZP;
    C* sP = (C*)&sZP[0][0][0][0];
    C* A = (C*)J.Y;
    const C* Vb = (const C*)J.Vbuf;
    const C* Zg = (const C*)J.Z;
    const int iv = i0 + 16 * wv + lr;
    const int rc = 2 * (t & 31);
    // row roles (t / 32 + 8 q) and their slots, fixed for the strip
    int rsl[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int i = i0 + (t >> 5) + 8 * q;
        rsl[q] = (i < n) ? gload_s(J.iprow + i) - k0 : -1;
    }
    // ---- tiles
    C pz[2][8];
    auto load_z = [&](int j0) {
        const int jz = j0 + 16 * wv + lr;
#pragma unroll
        for (int ch = 0; ch < 2; ++ch)
#pragma unroll
            for (int e = 0; e < 8; ++e)
                pz[ch][e] = gload(Zg + (jz < n && ch < nchunk ? (ch * PMAX + 8 * lk + e) * n + jz : 0));
    };
    f4v old[8];
    auto load_old = [&](int j0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = i0 + (t >> 5) + 8 * q, j = j0 + rc;
            const int sl = rsl[q];
            const bool zsrc = fused ? (sl >= 0 && sl < PMAX + nb2) : (sl >= 0 && sl < nb1);
            // element offsets fit in 32 bits (n <= MAX_LINE)
            const int off = (i < n && j < n) ? (zsrc ? sl * n + j : i * n + j) : 0;
            old[q] = gload2c((zsrc ? Zg : A) + off);
        }
    };
    // the first tile's loads are in flight while the V rows are split
    load_z(tj0 * 64);
    load_old(tj0 * 64);
    // ---- V rows of the strip (both chunks), split once
    {
        const int slv = (iv < n) ? gload_s(J.iprow + iv) - k0 : -1;
        f4v pv[2][4];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                pv[ch][u] = gload2c(Vb + (iv < n && ch < nchunk ? iv * VMAX + ch * PMAX + 8 * lk + 2 * u : 0));
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            const int nbc = (ch == 0) ? nb1 : nb2;
            // the first panel's pivot rows take no V1 part in a fused pair
            const bool okv = ch < nchunk && iv < n && !(fused && ch == 0 && slv >= 0 && slv < PMAX);
            float m = 0.f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const bool ok = okv && 8 * lk + q < nbc;
                pv[ch][q >> 1][2 * (q & 1)] = ok ? pv[ch][q >> 1][2 * (q & 1)] : 0.f;
                pv[ch][q >> 1][2 * (q & 1) + 1] = ok ? pv[ch][q >> 1][2 * (q & 1) + 1] : 0.f;
                m = fmaxf(m, fmaxf(fabsf(pv[ch][q >> 1][2 * (q & 1)]), fabsf(pv[ch][q >> 1][2 * (q & 1) + 1])));
            }
            m = fmaxf(m, __shfl_xor(m, 16, 64));
            m = fmaxf(m, __shfl_xor(m, 32, 64));
            const int e = split_exp(m);
            const float sc = exp2i(e);
            if (lk == 0) eV[ch][16 * wv + lr] = e;
            h8v rh, rl, ih, il;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                _Float16 h, l;
                split_f16(pv[ch][q >> 1][2 * (q & 1)] * sc, h, l);
                rh[q] = h;
                rl[q] = l;
                split_f16(pv[ch][q >> 1][2 * (q & 1) + 1] * sc, h, l);
                ih[q] = h;
                il[q] = l;
            }
            sV[ch][0][wv][lane] = rh;
            sV[ch][1][wv][lane] = rl;
            sV[ch][2][wv][lane] = ih;
            sV[ch][3][wv][lane] = il;
        }
    }
    for (int tl = 0; tl < ntl; ++tl) {
        const int j0 = (tj0 + tl) * 64;
        const int jz = j0 + 16 * wv + lr;
        __syncthreads();  // the previous tile's product reads are done (and the V split, first time)
        {
#pragma unroll
            for (int ch = 0; ch < 2; ++ch) {
                const int nbc = (ch == 0) ? nb1 : nb2;
                // Z1' : Z1 with the K2 columns zeroed
                const bool okz = ch < nchunk && jz < n && !(fused && ch == 0 && jz >= k2 && jz < kend);
                float m = 0.f;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const bool ok = okz && 8 * lk + q < nbc;
                    pz[ch][q].x = ok ? pz[ch][q].x : 0.f;
                    pz[ch][q].y = ok ? pz[ch][q].y : 0.f;
                    m = fmaxf(m, fmaxf(fabsf(pz[ch][q].x), fabsf(pz[ch][q].y)));
                }
                m = fmaxf(m, __shfl_xor(m, 16, 64));
                m = fmaxf(m, __shfl_xor(m, 32, 64));
                const int e = split_exp(m);
                const float sc = exp2i(e);
                if (lk == 0) eZ[ch][16 * wv + lr] = e;
                h8v rh, rl, ih, il;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    _Float16 h, l;
                    split_f16(pz[ch][q].x * sc, h, l);
                    rh[q] = h;
                    rl[q] = l;
                    split_f16(pz[ch][q].y * sc, h, l);
                    ih[q] = h;
                    il[q] = l;
                }
                sZ[ch][0][wv][lane] = rh;
                sZ[ch][1][wv][lane] = rl;
                sZ[ch][2][wv][lane] = ih;
                sZ[ch][3][wv][lane] = il;
            }
        }
        // the next tile's Z columns load while this tile is computed
        if (tl + 1 < ntl) load_z(j0 + 64);
        __syncthreads();
        // per chunk: re = Zr Vr + Zi (-Vi) and im = Zr Vi + Zi Vr accumulated
        // on MFMA (-Vi: the sign bits of the Vi fragments flipped), then
        // unscaled by the chunk's exponents into FP32 totals
        f4v tre[2][2], tim[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                tre[bi][bj] = f4v{0, 0, 0, 0};
                tim[bi][bj] = f4v{0, 0, 0, 0};
            }
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
            if (ch < nchunk) {
                f4v cre[2][2], cim[2][2];
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) {
                        cre[bi][bj] = f4v{0, 0, 0, 0};
                        cim[bi][bj] = f4v{0, 0, 0, 0};
                    }
                // terms outermost (hi hi, hi lo, lo hi): each term reads its
                // own fragments (re and im planes of the hi or lo part)
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const int xa = (r == 2) ? 1 : 0, yb = (r == 1) ? 1 : 0;
                    __builtin_amdgcn_sched_barrier(0);  // keep each term's reads beside its products
                    h8v a[2][2], b[2][3];  // [block][re, im]; b[.][2] = -Vi
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) {
                        a[bj][0] = sZ[ch][xa][wc * 2 + bj][lane];
                        a[bj][1] = sZ[ch][2 + xa][wc * 2 + bj][lane];
                    }
#pragma unroll
                    for (int bi = 0; bi < 2; ++bi) {
                        b[bi][0] = sV[ch][yb][wr * 2 + bi][lane];
                        b[bi][1] = sV[ch][2 + yb][wr * 2 + bi][lane];
                        b[bi][2] = -b[bi][1];
                    }
#define H3_MM(ACC, XP, YP)                                                                          \
    _Pragma("unroll") for (int bi = 0; bi < 2; ++bi) _Pragma("unroll") for (int bj = 0; bj < 2; ++bj) \
        ACC[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[bj][XP], b[bi][YP], ACC[bi][bj], 0, 0, 0);
                    H3_MM(cre, 0, 0)  // Zr Vr
                    H3_MM(cim, 0, 1)  // Zr Vi
                    H3_MM(cre, 1, 2)  // Zi (-Vi)
                    H3_MM(cim, 1, 0)  // Zi Vr
#undef H3_MM
                }
#pragma unroll
                for (int bi = 0; bi < 2; ++bi) {
                    const int ei = eV[ch][wr * 32 + bi * 16 + lr];
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const float sc = exp2i(-(ei + eZ[ch][wc * 32 + bj * 16 + 4 * lk + v]));
                            tre[bi][bj][v] += cre[bi][bj][v] * sc;
                            tim[bi][bj][v] += cim[bi][bj][v] * sc;
                        }
                }
            }
        }
        __syncthreads();  // every wave is done with the Z planes
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const int o = (wr * 32 + bi * 16 + lr) * H3_PROW + wc * 32 + bj * 16 + 4 * lk;
                *(f4v*)(sP + o) = f4v{tre[bi][bj][0], tim[bi][bj][0], tre[bi][bj][1], tim[bi][bj][1]};
                *(f4v*)(sP + o + 2) = f4v{tre[bi][bj][2], tim[bi][bj][2], tre[bi][bj][3], tim[bi][bj][3]};
            }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = (t >> 5) + 8 * q, i = i0 + r, j = j0 + rc;
            if (i < n && j < n) {
                const int sl = rsl[q];
                const bool p1 = fused && sl >= 0 && sl < PMAX;
                // rows replaced by Z: the panel's pivots (single) or the
                // second panel's pivots (fused)
                const bool zrow = fused ? (sl >= PMAX && sl < PMAX + nb2) : (sl >= 0 && sl < nb1);
                const f4v pp = *(const f4v*)(sP + r * H3_PROW + rc);
                f4v o = old[q];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int jj = j + h;
                    const bool ex = p1 ? (jj >= k2 && jj < kend) : (jj >= k0 && jj < kend);
                    const float orr = ex ? 0.f : old[q][2 * h], oii = ex ? 0.f : old[q][2 * h + 1];
                    o[2 * h] = zrow ? old[q][2 * h] : orr - pp[2 * h];
                    o[2 * h + 1] = zrow ? old[q][2 * h + 1] : oii - pp[2 * h + 1];
                }
                gstore2c(A + (i * n + j), o);
            }
        }
        if (tl + 1 < ntl) load_old(j0 + 64);
    }
}

void launch_trailing_h3(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb1,
                        int nb2) {
    const int nts = (n + 63) / 64, ngrp = (nts + H3_TPW - 1) / H3_TPW;
    const int total = nts * ngrp * njobs;
    k_trailing_h3s<H3_TPW><<<total, 256, 0, s>>>(jobs_d, k0, nb1, nb2, nts, ngrp, total);
}


template <typename R>
void launch_trailing(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb1,
                     int nb2) {
    const int ntx = (n + 63) / 64;
    if (std::is_same<R, float>::value && n % 4 == 0) {
        const int total = ntx * ntx * njobs;
        k_trailing_f32<<<total, 256, 0, s>>>(jobs_d, k0, nb1, nb2, ntx, total);
        return;
    }
    dim3 grid(ntx, ntx, njobs);
    k_trailing<R><<<grid, 256, 0, s>>>(jobs_d, k0, nb1, nb2);
}
template void launch_trailing<double>(hipStream_t, const LineJob*, int, int, int, int, int);
template void launch_trailing<float>(hipStream_t, const LineJob*, int, int, int, int, int);


}  // namespace fdfd
