// Internal declarations shared by the HIP kernels and the host runtime of
// libfdfd.  Complex values are interleaved (re, im) pairs: double2 for FP64,
// float2 for FP32.  Field-shaped arrays use the natural layout of the
// reference (x fastest: index ix + iy*Nx).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace fdfd {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
template <typename R> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename R> using cplx = typename cplx_t<R>::type;

template <typename C> __host__ __device__ inline C mk(decltype(C{}.x) re, decltype(C{}.x) im) {
    C z; z.x = re; z.y = im; return z;
}
template <typename C> __host__ __device__ inline C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __host__ __device__ inline C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __host__ __device__ inline C cmul(C a, C b) {
    return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a*conj(b)
template <typename C> __host__ __device__ inline C cmulc(C a, C b) {
    return mk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a)*b
template <typename C> __host__ __device__ inline C cconjmul(C a, C b) {
    return mk<C>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename C> __host__ __device__ inline C cneg(C a) { return mk<C>(-a.x, -a.y); }
// component-wise select (a struct-valued ?: can be lowered to a stack array)
template <typename C> __host__ __device__ inline C csel(bool p, C a, C b) {
    return mk<C>(p ? a.x : b.x, p ? a.y : b.y);
}
template <typename C> __host__ __device__ inline C cscale(C a, decltype(C{}.x) s) { return mk<C>(a.x * s, a.y * s); }
// 1/z with scaling against overflow (Smith's method)
template <typename C> __host__ __device__ inline C crcp(C z) {
    using R = decltype(C{}.x);
    R ar = z.x < 0 ? -z.x : z.x, ai = z.y < 0 ? -z.y : z.y;
    if (ar >= ai) {
        R t = z.y / z.x, d = z.x + z.y * t;
        return mk<C>(R(1) / d, -t / d);
    } else {
        R t = z.x / z.y, d = z.y + z.x * t;
        return mk<C>(t / d, R(-1) / d);
    }
}
template <typename C> __host__ __device__ inline C cdiv(C a, C b) { return cmul(a, crcp(b)); }
// LAPACK-style pivot magnitude |re| + |im| (izamax / cabs1).  The sign bits
// are cleared, so a signed zero entry has magnitude +0 (a comparison-based
// |x| returns -0 for -0, whose bit pattern would order above every positive
// magnitude in an integer argmax).
__host__ __device__ inline float rabs(float x) { return __builtin_fabsf(x); }
__host__ __device__ inline double rabs(double x) { return __builtin_fabs(x); }
template <typename C> __host__ __device__ inline decltype(C{}.x) cabs1(C z) {
    return rabs(z.x) + rabs(z.y);
}
template <typename C> __host__ __device__ inline decltype(C{}.x) cnorm2(C z) { return z.x * z.x + z.y * z.y; }

__host__ __device__ inline float2 to_f(double2 a) { return make_float2((float)a.x, (float)a.y); }
__host__ __device__ inline double2 to_d(float2 a) { return make_double2((double)a.x, (double)a.y); }
__host__ __device__ inline double2 to_d(double2 a) { return a; }
template <typename C> __host__ __device__ inline C from_d(double2 a);
template <> __host__ __device__ inline double2 from_d<double2>(double2 a) { return a; }
template <> __host__ __device__ inline float2 from_d<float2>(double2 a) { return to_f(a); }

// ---------------------------------------------------------------------------
// physical constants: src/types.jl:6-9 (SI, as in the reference)
// ---------------------------------------------------------------------------
constexpr double EPS0 = 8.85418782e-12;
constexpr double MU0 = 1.25663706e-6;
constexpr double PML_M = 3.5;    // src/pml.jl:1
constexpr double PML_LNR = -12;  // src/pml.jl:1

// ---------------------------------------------------------------------------
// operator coefficients.  The engine works with the scaled operator
// A' = sigma_s A (sigma_s = mu0' dx^2 for TM, eps0' dx^2 for TE) so that the
// entries are O(1); a 5-point periodic stencil
//   (A'x)_ij = cW x_{i-1,j} + cE x_{i+1,j} + cS x_{i,j-1} + cN x_{i,j+1} + cC x_ij
// with the five coefficient planes stored contiguously (plane k at coef+k*N).
// ---------------------------------------------------------------------------
enum CoefPlane { CW = 0, CE = 1, CS = 2, CN = 3, CC = 4 };

struct GridDesc {
    int Nx, Ny;
    double dx, dy;
    int npx, npy;
    double L0;
};

// One line-block inversion in flight (Gauss-Jordan with partial pivoting,
// in place, pivot rows tracked by index instead of swapped).
struct LineJob {
    void* Y;          // n x n row-major complex (R); input D, output Y
    int* used;        // n: row already chosen as a pivot
    int* prow;        // n: prow[c] = pivot row of column c
    int* iprow;       // n: inverse of prow (-1 until pivoted)
    void* Vbuf;       // n x VMAX panel copies V (two panels of a fused pair)
    void* LU;         // 2 x PMAX x PMAX: V[piv, :]^-1 of the pivot block, then U^-1
    int* piv;         // VMAX: pivot rows of the two panels
    void* Z;          // VMAX x n: V[piv, :]^-1 * [A(piv, :) | I] of the two panels
    int* win;         // 64: candidate-row window (first unpivoted rows by index)
    int* verr;        // set to 1 when a row outside the window beats a chosen pivot
    double* pivstat;  // [0] = min |pivot|, [1] = max |pivot|, [2] = zero-pivot flag
    int n;
};

// ---------------------------------------------------------------------------
// Block cyclic reduction (lines on the y-periodic ring).  All blocks are
// n x n row-major complex (R).  A line inverted by Gauss-Jordan keeps its
// result Y in place with X = D^-1 read as X[a, b] = Y[prow[a], iprow[b]].
// ---------------------------------------------------------------------------
// Dense block formed from diagonal couplings (first reduction level):
//   out[a, b] = own(a, b) + [a == b] (dd0 + dd1)[a]
//               - sum_t (rs_t0 + rs_t1)[a] X_t[a, b] (cs_t0 + cs_t1)[b]
// own = the line's periodic tridiagonal operator (sub, diag, sup; nullable).
// Row vectors are indexed with the output line's stride si, column vectors
// with the same stride (every line has the same layout).
struct FormJob {
    void* out;
    const double2 *sub, *diag, *sup;
    const double2* dd[2];
    const void* Y[2];
    const int* prow[2];
    const int* iprow[2];
    const double2* rs[2][2];
    const double2* cs[2][2];
    int nt;
    long si;
};

// Inverse of a line's own periodic tridiagonal block (first reduction level)
// by Givens QR: T[a][a] = diag + dd0 + dd1, T[a][a+1] = sup, T[a][a-1] = sub
// (indices wrap).  out receives X = T^-1 in natural order (prow = iprow =
// identity); work holds the rotations and R (FP64).
struct TriJob {
    void* out;
    const double2 *sub, *diag, *sup;
    const double2* dd[2];
    int* prow;
    int* iprow;
    double* pivstat;
    double* work;
    long si;
};

// First-level line solve of the reduced solve: out = T^-1 in along one line
// (cells at stride si), T the line's periodic tridiagonal block, by its
// Givens QR work w (tri_qr.h layout, FP64).
struct TrsJob {
    const double2* in;
    double2* out;
    const double* w;
    long si;
};

// Split operand of the three-product GEMM: the four fp16 planes (re hi,
// re lo, im hi, im lo) of an n x n complex64 block, as is (row r, k =
// column) or transposed with a row gather (row c, k: src[map[k]][c]), each
// row r scaled by 2^e_r (e_r in exps, chosen so the row's largest |re|, |im|
// is at most 2^14), in the tiled operand layout: tiles of 16 rows x 32 k,
// tile (rb, kb) at (rb * nkb + kb) * 512, entry (r, k) of a tile at
// (k / 8) * 128 + r * 8 + k % 8, so that lane l of an MFMA fragment reads the
// 16 bytes at 16 l.  Rows and k past n are zero.
inline __host__ __device__ int b3_nrb(int n) { return (n + 15) / 16; }
inline __host__ __device__ int b3_nkb(int n) { return (n + 31) / 32; }
inline __host__ __device__ long b3_plane(int n) { return (long)b3_nrb(n) * b3_nkb(n) * 512; }
// bytes of one split block: four planes of fp16, then the row exponents
// (ints), padded to 256 bytes
inline __host__ __device__ size_t b3_block_bytes(int n) {
    const size_t b = 4 * (size_t)b3_plane(n) * 2 + (size_t)b3_nrb(n) * 16 * sizeof(int);
    return (b + 255) / 256 * 256;
}
struct SplitJob {
    const void* src;
    void* dst;
    const int* map;
    int* exps;  // b3_nrb(n) * 16 row exponents
    int trans;
};

// out = [acc ? out : 0] -/+ sum_t A_t * B_t[bmap_t, :]   (B rows gathered;
// sub = 1 subtracts the products, sub = 0 adds them)
struct GemmJob {
    void* out;
    const void* A[2];
    const void* B[2];
    const int* bmap[2];
    const int* ea[2];  // split operands: row exponents of A_t and of B_t^T
    const int* eb[2];
    int nt;
    int acc;
    int sub;
};

// Dense matrix-vector step of the reduced solve (vectors are field-shaped,
// a line's entry i at x[i * si]):
//   z = sum_t M_t w_t,  w_t[c] = x_t[imap_t ? imap_t[c] : c]
//   out[omap ? omap[r] : r] = base ? base[r] - z[r] : z[r]
struct MvJob {
    const void* M[2];
    const int* imap[2];
    const double2* x[2];
    const double2* base;
    double2* out;
    const int* omap;
    int nt;
    long si;
};

// Diagonal step of the reduced solve (first level):
//   out[i] = x0[i] - sum_t (d_t0 + d_t1)[i] x_t[i]
struct DiagJob {
    double2* out;
    const double2* x0;
    const double2* x[2];
    const double2* d[2][2];
    int nt;
    long si;
};

// One 1-D eigenmode problem (src/solver/eigen.jl:6-29): the periodic
// tridiagonal operator of the slice, shift sigma, Arnoldi storage and outputs.
struct ModeJob {
    const double2* eps;  // n relative permittivities along the slice
    double2* V;          // (ncv + 1) * n Arnoldi basis
    double2* W;          // ncv * n restart workspace
    double* trw;         // factors of A - sigma I (tri::work_doubles(n))
    double2* vec;        // out: nev eigenvectors, vector k at vec + k * n
    double2* beta;       // out: nev propagation constants
    int* info;           // out: status, converged pairs, restarts, OP applications
    double2 sigma;       // (omega / c0' * neff)^2
    double k0sq;         // omega^2 mu0' eps0'
    double idx2;         // 1 / dx^2 of the 1-D grid
    int n, pol;
    int ncv;             // Arnoldi basis size (nev < ncv <= min(n, MODE_NCV))
};
// status values of ModeJob::info[0]
enum { MODE_OK = 0, MODE_SINGULAR = 1, MODE_EIG_FAIL = 2 };
// largest Arnoldi basis of the mode kernel (its projected matrices live in LDS)
constexpr int MODE_NCV = 32;

// ---------------------------------------------------------------------------
// launch wrappers (implemented in the .hip files)
// ---------------------------------------------------------------------------
void launch_sfactor(hipStream_t s, double2* sxf, double2* sxb, double2* syf, double2* syb,
                    const GridDesc& g, double omega, double eps0n, double eta0);
void launch_coeffs(hipStream_t s, double2* coef, const double2* eps, const double2* sxf,
                   const double2* sxb, const double2* syf, const double2* syb,
                   const GridDesc& g, int pol, double kdiag, int form);
// y[j] += sign * c[j] * (conj(d) x[j+1] + d x[j-1]) for the nf sideband
// blocks of N cells (modulation.jl:95-99); c has nf entries
void launch_mf_couple(hipStream_t s, double2* y, const double2* x, const double2* d,
                      const double* c, int nf, long N, double sign);
// full-vector mode operator (unknowns Hx, Hy): planes of s (A_bb - sigma),
// the off-diagonal coupling y += cs (A_xy x_y ; A_yx x_x), and the six fields
void launch_vmode_coeffs(hipStream_t st, double2* coef, const double2* eps, const GridDesc& g,
                         double k0sq, double2 sig, double s, int block);
void launch_vmode_offdiag(hipStream_t st, double2* y, const double2* x, const double2* eps,
                          const GridDesc& g, double cs);
void launch_vmode_fields(hipStream_t st, double2* out, const double2* h, const double2* eps,
                         const GridDesc& g, double2 beta, double omega, double eps0n);
// Poynting maps (S: Sx then Sy, N doubles each) and plane fluxes (int4 =
// normal, index, lo, hi) of a 3 x N field array
void launch_poynting(hipStream_t s, double* S, const double2* f, int Nx, int Ny, int pol);
void launch_flux(hipStream_t s, double* out, const double2* f, const int4* planes, int nplanes,
                 int Nx, int Ny, int pol, double dx, double dy);
// matrix-free stencil of a driven (form 0, TM or TE) or sideband (form 1,
// TM) operator: weights from the 1-D inverse stretch vectors and eps
struct StencilOp {
    int pol, form, Nx, Ny;
    const double2 *eps, *sxf, *sxb, *syf, *syb;
    double ry, kdiag;
};
void launch_stencil(hipStream_t s, double2* y, const double2* x, const StencilOp& op,
                    int transpose, const double2* b, long rs = 0, const int* ridx = nullptr,
                    int na = 1);
// several right-hand sides per launch (k_multi.hip): sources ridx[0..na) of
// buffers whose source r starts at r * rs; the matrix-vector step runs on the
// matrix cores (16 sources per chunk, inputs packed into `pack`, at least
// mv_pack_bytes(njobs, n, fp64) bytes) when mfma is set, else on the vector ALUs
template <typename R> int mv_multi_prepare(int n);
size_t mv_pack_bytes(int njobs, int n, int fp64);
template <typename R>
void launch_mv_multi(hipStream_t s, const MvJob* jobs_d, int njobs, int n, long rs, const int* ridx,
                     int na, bool mfma, void* pack);
void launch_diag_multi(hipStream_t s, const DiagJob* jobs_d, int njobs, int n, long rs,
                       const int* ridx, int na);
int trisolve_multi_prepare(int n);
void launch_trisolve_multi(hipStream_t s, const TrsJob* jobs_d, int njobs, int n, long rs,
                           const int* ridx, int na);
void launch_multidot_b(hipStream_t s, const double2* V, long ldv, long rs, int k, const double2* w,
                       long wrs, long n, double2* partial, double2* out, int nblk,
                       const int* ridx, int na);
void launch_gemv_update_b(hipStream_t s, double2* w, long wrs, const double2* V, long ldv, long rs,
                          const int* kcnt, const double2* h, int kstride, double sign, long n,
                          const int* ridx, int na);
void launch_norm2_b(hipStream_t s, const double2* x, long rs, long n, double* partial, double* out,
                    int nblk, const int* ridx, int na);
void launch_scale_b(hipStream_t s, double2* y, long yrs, const double2* x, long xrs, const double2* a,
                    long n, const int* ridx, int na);
void launch_apply(hipStream_t s, double2* y, const double2* x, const double2* coef,
                  int Nx, int Ny, int transpose, const double2* b /*nullable: r=b-Ax*/);
void launch_scale_copy(hipStream_t s, double2* y, const double2* x, double2 a, long n);
void launch_zero(hipStream_t s, double2* y, long n);
void launch_multidot(hipStream_t s, const double2* V, long ldv, int k, const double2* w,
                     long n, double2* partial, int nblk);
void launch_reduce_partials(hipStream_t s, const double2* partial, int nblk, int k, double2* out);
void launch_gemv_update(hipStream_t s, double2* w, const double2* V, long ldv, int k,
                        const double2* h, double sign, long n);
void launch_norm2(hipStream_t s, const double2* x, long n, double* partial, int nblk);
void launch_reduce_real(hipStream_t s, const double* partial, int nblk, double* out);
void launch_fields(hipStream_t s, double2* out, const double2* x, const double2* eps,
                   const double2* sxf, const double2* sxb, const double2* syf,
                   const double2* syb, const GridDesc& g, int pol, double2 w,
                   int eigen_mode);
void launch_mul_eps(hipStream_t s, double2* y, const double2* x, const double2* eps,
                    double2 a, long n);
void launch_grad(hipStream_t s, double* grad, const double2* e, const double2* lam,
                 double2 coef, long n);
void launch_field_adjoint(hipStream_t s, double2* out, const double2* dfield, const double2* sxb,
                          const double2* syb, const GridDesc& g, double2 c, double scale);

template <typename R>
void launch_panel(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb);
template <typename R>
void launch_wpanel(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb,
                   int fresh, int second);
template <typename R>
void launch_zbuild(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb,
                   int fresh, int verify, int second, int k1, int nb1);
template <typename R>
void launch_trailing(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb1,
                     int nb2);
// mixed path, n % 4 == 0: the trailing update on fp16x3 products
void launch_trailing_h3(hipStream_t s, const LineJob* jobs_d, int njobs, int n, int k0, int nb1,
                        int nb2);
template <typename R>
void launch_form(hipStream_t s, const FormJob* jobs_d, int njobs, int n);
void launch_gjreset(hipStream_t s, const LineJob* jobs_d, int njobs, int n);
void launch_trifactor(hipStream_t s, const TriJob* jobs_d, int njobs, int n);
template <typename R>
void launch_triinv(hipStream_t s, const TriJob* jobs_d, int njobs, int n);
template <typename R>
void launch_gemm(hipStream_t s, const GemmJob* jobs_d, int njobs, int n, int gauss3, int waves);
void launch_gemm_b3(hipStream_t s, const GemmJob* jobs_d, int njobs, int n);
void launch_split(hipStream_t s, const SplitJob* jobs_d, int njobs, int n);
template <typename R>
void launch_mv(hipStream_t s, const MvJob* jobs_d, int njobs, int n);
void launch_diag(hipStream_t s, const DiagJob* jobs_d, int njobs, int n);
// k_trisolve keeps the line's vector and a double-buffered chunk of its QR
// work in LDS (16 n + 20480 bytes); trisolve_prepare sets the kernel's
// dynamic LDS limit (call before any launch or capture; returns a HIP error
// code)
int trisolve_prepare(int n);
void launch_trisolve(hipStream_t s, const TrsJob* jobs_d, int njobs, int n);
void launch_wscale(hipStream_t s, double2* out, const double2* in, const double2* sxf,
                   const double2* syf, int Nx, int Ny, int inverse);
void launch_mode(hipStream_t s, const ModeJob* jobs_d, int njobs, int nev, double tol,
                 int maxiter);

// largest line-block size the panel kernel supports (one row per thread)
constexpr int MAX_LINE = 1024;
// widest Gauss-Jordan panel (L, U, pivot-block inverse are PMAX x PMAX)
constexpr int PMAX = 32;
// two panels are applied in one trailing pass: the panel copy V is n x VMAX,
// Z is VMAX x n, pivots VMAX (first panel at 0, second at PMAX)
constexpr int VMAX = 2 * PMAX;
// rows in the pivot-candidate window of the fast panel (one wave)
constexpr int WIN = 64;
// threshold of the windowed pivot search: a window pivot is accepted when no
// unpivoted row outside the window is larger than 1 / PIV_U times it (the
// threshold partial pivoting of sparse LU codes; 1 would be strict GEPP)
constexpr double PIV_U = 0.1;

}  // namespace fdfd
