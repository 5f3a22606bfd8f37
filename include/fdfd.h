/*
 * fdfd.h -- C ABI of libfdfd, the MI355X (gfx950) engine for the 2-D FDFD
 * driven solve and shift-invert eigenfrequency path of FDFD.jl.
 *
 * Plain C: no C++ or HIP types cross this boundary.  Complex numbers are
 * interleaved (re, im) doubles, identical in memory to Julia ComplexF64,
 * C `double _Complex` and numpy complex128.  2-D arrays are Nx x Ny with x
 * fastest (Julia column-major, linear index ix + iy*Nx, 0-based).
 *
 * Every function returns an int status: 0 = FDFD_OK, negative = error; the
 * message of the last error on a handle is returned by fdfd_last_error().
 * A handle owns one HIP stream and every device buffer it allocates; it is
 * not thread-safe (distinct handles are independent).  Host buffers passed
 * in are borrowed for the duration of the call only.  Functions whose names
 * end in `_d` take device pointers (hipMalloc'd memory on the handle's
 * device) instead of host pointers and do not copy over PCIe.
 *
 * Reference interfaces replaced (FDFD.jl, Julia):
 *   solve(d::Device, pol)                 src/solver/driven.jl:4-59
 *   dolinearsolve(A, b, sym)  (lu(A)\b)   src/solver/solver.jl:4-41
 *   eigenfrequency(d, pol, nev; which)    src/solver/eigen.jl:69-115
 *   eigenmode(d::Device{1}, pol, neff, n) src/solver/eigen.jl:6-29
 *   eigenmode(d::AbstractDevice{2}, ...)  src/solver/eigen.jl:31-66
 *   solve(d::ModulatedDevice)             src/solver/modulation.jl:35-119
 *   poynting / flux_surface_integral      src/flux.jl:15-66
 *     (get_modes / setup_mode!)           src/device.jl:118-149
 *   S_create / create_sfactor             src/pml.jl:1-63
 *   delta / grid_average                  src/grid.jl:128-162
 *   FieldTM / FieldTE layout              src/data.jl:50-78
 */
#ifndef FDFD_H
#define FDFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdfd_handle fdfd_handle;

typedef struct {
    double re;
    double im;
} fdfd_complex;

/* Polarization: src/types.jl:20 (@enum Polarization TM=1 TE=2). */
enum { FDFD_TM = 1, FDFD_TE = 2 };

/* Factor precision.  FDFD_FP64: complex128 block factors (the plain path).
 * FDFD_FP32_IR: complex64 block factors, accuracy restored by FP64 flexible
 * GMRES on the exact FP64 operator.  The complex64 products run by default as
 * fp16x3 MFMA products of row-scaled split operands (about 22-bit products,
 * FP32 accumulation); the environment switches FDFD_MIXED_GEMM=fp32 and
 * FDFD_MIXED_TRAIL=fp32, read at fdfd_create, select FP32 MFMA products for
 * the Schur GEMMs and the Gauss-Jordan trailing updates instead. */
enum { FDFD_FP64 = 0, FDFD_FP32_IR = 1 };

/* Eigenvalue selection for fdfd_eigenfrequency (Arpack `which`, applied to
 * the eigenvalues theta = 1/(lambda - sigma) of the shift-inverted operator):
 * LM = largest |theta| = nearest the shift; LR = largest Re(theta). */
enum { FDFD_WHICH_LM = 0, FDFD_WHICH_LR = 1 };

/* Status codes. */
enum {
    FDFD_OK = 0,
    FDFD_ERR_INVALID = -1,   /* bad argument / unsupported size            */
    FDFD_ERR_OOM = -2,       /* device allocation failed                    */
    FDFD_ERR_SINGULAR = -3,  /* zero pivot in a line block (stats.line)     */
    FDFD_ERR_NOCONV = -4,    /* refinement / Arnoldi did not converge       */
    FDFD_ERR_HIP = -5,       /* HIP runtime error                           */
    FDFD_ERR_NOTFOUND = -6,  /* unknown factor id                           */
    FDFD_ERR_STATE = -7      /* call order violated (e.g. no grid set)      */
};

typedef struct {
    int iterations;        /* GMRES iterations (all restarts, all RHS)     */
    int restarts;
    double residual;       /* final max relative residual ||b-Ax||/||b||    */
    double t_factor_ms;    /* device time of the block factorization        */
    double t_solve_ms;     /* device time of the refined solve              */
    double pivot_growth;   /* max |pivot| / min |pivot| over all line blocks */
    int singular_line;     /* line index of a zero pivot, -1 if none        */
    int nconv;             /* converged eigenpairs (eigenfrequency)         */
    int pivot_fallback;    /* 1: a pivot outside the candidate window won a
                              step and the factorization was redone with the
                              full-column pivot search                       */
} fdfd_stats;

/* ---- library / handle --------------------------------------------------- */
int fdfd_version(int* major, int* minor, int* patch);
const char* fdfd_last_error(const fdfd_handle* h);
/* Which kernel variant a solve used ("gfx950"), for provenance checks. */
const char* fdfd_arch(void);

int fdfd_create(int device_ordinal, fdfd_handle** out);
int fdfd_destroy(fdfd_handle* h);

/* Grid(dh, Npml, xrange, yrange; L0): src/grid.jl:26-33.  dx, dy are the
 * cell sizes span/N (src/grid.jl:68-74) in units of L0; PML m=3.5 and
 * lnR=-12 are fixed as in src/pml.jl:1,35-38. */
int fdfd_set_grid(fdfd_handle* h, int Nx, int Ny, double dx, double dy,
                  int npml_x, int npml_y, double L0);

/* Relative permittivity Device.eps_r (src/device.jl:21), Nx*Ny complex. */
int fdfd_set_eps(fdfd_handle* h, const fdfd_complex* eps);
int fdfd_set_eps_d(fdfd_handle* h, const void* d_eps);

/* ---- factorization ------------------------------------------------------ */
/* Assembles the TM (driven.jl:35) or TE (driven.jl:45) operator at angular
 * frequency omega (rad/s, normalised units c0/L0 as in device.jl:40) on the
 * device and factors it by block cyclic reduction of its ring of grid lines
 * (the y-periodic corner blocks close the ring): every reduction level
 * inverts every other line's dense block by Gauss-Jordan elimination with
 * threshold partial pivoting (|pivot| >= 0.1 x the column's largest
 * candidate) and folds its couplings into its neighbours with complex GEMMs.
 * The factorization is a direct solver of A.  Replaces lu(A) of
 * solver.jl:35. */
int fdfd_factor(fdfd_handle* h, int pol, double omega, int precision,
                int* factor_id, fdfd_stats* stats);
int fdfd_release_factor(fdfd_handle* h, int factor_id);

/* ---- solves ------------------------------------------------------------- */
/* Driven solve for nrhs source arrays src (Nx*Ny*nrhs, Device.src
 * src/device.jl:22): b = 1im*omega*src (driven.jl:36,46), A x = b solved to
 * relative residual <= tol, then field recovery (driven.jl:40-41, 50-51).
 * out is Nx*Ny*3*nrhs: per RHS the three components in FieldTM order
 * (Ez, Hx, Hy) or FieldTE order (Hz, Ex, Ey), x fastest, component slowest
 * -- the memory layout of cat(reshape.(...), dims=3) in data.jl:62,77. */
int fdfd_solve(fdfd_handle* h, int factor_id, int nrhs, const fdfd_complex* src,
               fdfd_complex* out, double tol, fdfd_stats* stats);
int fdfd_solve_d(fdfd_handle* h, int factor_id, int nrhs, const void* d_src,
                 void* d_out, double tol, fdfd_stats* stats);

/* Plain linear solves with the reference-scaled operator A of driven.jl:35/45:
 * x = A \ b  (transpose=0) or x = A^T \ b (transpose=1).  b, x: Nx*Ny*nrhs. */
int fdfd_linsolve(fdfd_handle* h, int factor_id, int transpose, int nrhs,
                  const fdfd_complex* b, fdfd_complex* x, double tol,
                  fdfd_stats* stats);

/* y = A x or y = A^T x with the exact (periodic, PML-stretched) operator. */
int fdfd_apply(fdfd_handle* h, int factor_id, int transpose,
               const fdfd_complex* x, fdfd_complex* y);

/* rel = ||b - A x||_2 / ||b||_2 in FP64 on the device. */
int fdfd_residual(fdfd_handle* h, int factor_id, const fdfd_complex* x,
                  const fdfd_complex* b, double* rel);

/* Adjoint gradient with respect to eps_r (TM): for A(eps) e = b and an
 * adjoint field lam solving A^T lam = dJ/de,
 *   grad[k] = -Re(weight * lam[k] * omega^2 * eps0' * e[k])
 * (dA/deps_k = omega^2 eps0' e_k e_k^T from driven.jl:21,35).  For
 * J = |m^T e|^2 pass weight = 2*conj(m^T e) and lam = A^-T m. */
int fdfd_adjoint_grad(fdfd_handle* h, int factor_id, const fdfd_complex* e_fwd,
                      const fdfd_complex* lam, fdfd_complex weight, double* grad);

/* Pullback (reverse-mode derivative) of a TM driven solve with respect to
 * eps_r: the body of the ChainRulesCore rrule of solve(d::Device, TM)
 * (driven.jl:4-59; the reference has no derivative).  For the fields
 * F = (Ez, Hx, Hy) of fdfd_solve on factor_id and a real loss L(F) whose
 * cotangent is dfields = dL/dRe F + i dL/dIm F (Nx*Ny*3, the fdfd_solve
 * layout), it returns grad[k] = dL/d eps_r[k] (Nx*Ny doubles):
 *   bar e = bar Ez + Mx^H bar Hx + My^H bar Hy   (H = M Ez, driven.jl:40-41)
 *   A^T lam = conj(bar e)                       (FGMRES on A^T, the factor's
 *                                                transposed solve, to tol)
 *   grad[k] = -Re(lam[k] w^2 eps0' Ez[k])       (dA/deps_k, driven.jl:21,35)
 * ez is the Ez component of the forward solve (Nx*Ny).  The source is taken
 * as independent of eps_r.  FDFD_ERR_INVALID for TE factors. */
int fdfd_solve_pullback(fdfd_handle* h, int factor_id, const fdfd_complex* ez,
                        const fdfd_complex* dfields, double tol, double* grad,
                        fdfd_stats* stats);
int fdfd_solve_pullback_d(fdfd_handle* h, int factor_id, const void* d_ez,
                          const void* d_dfields, double tol, void* d_grad,
                          fdfd_stats* stats);

/* ---- eigenfrequencies --------------------------------------------------- */
/* Shift-invert Arnoldi (eigen.jl:69-115): eigenvalues lambda of the TM
 * operator diag(1/eps) (dxf dxb + dyf dyb) nearest sigma = -w0^2 mu0' eps0'
 * (TE: dxf Tex^-1 dxb + dyf Tey^-1 dyb, sigma = -w0^2 mu0'), returned as
 * omega = sqrt(-lambda/(mu0' eps0')) (TE: sqrt(-lambda/mu0')), in the order
 * of `which`: nearest sigma first (LM) or largest Re 1/(lambda - sigma) first
 * (LR).  precision as for fdfd_factor; other values of precision or which
 * are FDFD_ERR_INVALID.  fields (may be NULL): nev * Nx*Ny*3 eigenfields with
 * the forward-difference H recovery of eigen.jl:90-91 / 108-109, each
 * eigenvector normalised to unit 2-norm.  The shifted operator is factored
 * once and the factors are reused for every Arnoldi step. */
int fdfd_eigenfrequency(fdfd_handle* h, int pol, double omega0, int nev,
                        int ncv, int which, int precision, double tol,
                        int maxiter, fdfd_complex* omegas, fdfd_complex* fields,
                        fdfd_stats* stats);

/* ---- 1-D eigenmodes (mode sources) -------------------------------------- */
/* One slice of eps, as get_modes builds it (src/device.jl:124-149): n cells of
 * size dh along the slice, periodic, no PML. */
typedef struct {
    int n;                    /* cells of the slice, 3 .. 16384                  */
    int pol;                  /* FDFD_TM or FDFD_TE                              */
    double omega;             /* angular frequency (rad/s)                        */
    double dh;                /* cell size along the slice, in units of L0        */
    fdfd_complex neff;        /* effective-index guess                            */
    const fdfd_complex* eps;  /* n relative permittivities along the slice (host) */
} fdfd_mode_problem;

/* eigenmode(d::Device{1}, pol, neff, nev) (src/solver/eigen.jl:6-29) for
 * nprob slices at once, one workgroup each.  The 1-D grid is
 * Grid(n, [0 n*dh]) (dx = (n*dh)/n); the operator is
 *   TM: omega^2 mu0' eps0' diag(eps) + dxf dxb
 *   TE: omega^2 mu0' eps0' diag(eps) + diag(eps) dxf diag(1/avg(eps)) dxb
 * with avg = grid_average (src/grid.jl:157-158).  Shift-invert Arnoldi with
 * sigma = (omega/c0' neff)^2 returns the nev eigenvalues beta^2 nearest
 * sigma, nearest first, as beta = sqrt(beta^2) (principal branch).
 *   ncv <= 0: max(20, 2 nev + 1) (the Arpack.jl default), at most 32, and
 *             clamped to n per slice; nev < min(ncv, n) is required.
 *   L0 <= 0: 1e-6.  tol <= 0: 1e-12 (Ritz residual relative to the Ritz
 *             value of OP).  maxiter <= 0: 300 restarts.
 *   beta:    nprob * nev, problem p at p * nev.
 *   vectors: may be NULL; problem p's nev vectors start at nev * (n_0 + ..
 *            + n_{p-1}), vector k at + k * n_p; unit 2-norm, largest entry
 *            real and positive (ARPACK's phase is arbitrary; setup_mode! uses
 *            abs.(v) only).
 *   nconv:   may be NULL; nprob converged-pair counts.
 * FDFD_ERR_SINGULAR if sigma is an eigenvalue of a slice's operator
 * (stats.singular_line = that problem), FDFD_ERR_NOCONV if fewer than nev
 * pairs converged (outputs still written). */
int fdfd_eigenmode(fdfd_handle* h, int nprob, const fdfd_mode_problem* probs,
                   int nev, int ncv, double L0, double tol, int maxiter,
                   fdfd_complex* beta, fdfd_complex* vectors, int* nconv,
                   fdfd_stats* stats);

/* ---- full-vector waveguide modes --------------------------------------- */
/* eigenmode(d::AbstractDevice{2}, neff, nev) (src/solver/eigen.jl:31-66) on
 * the handle's grid and eps (periodic differences, no PML, as the reference):
 * the nev eigenvalues beta^2 of the (Hx, Hy) operator of eigen.jl:45 nearest
 * (omega/c0' neff)^2, nearest first, as beta = sqrt(beta^2).
 *   fields: may be NULL; nev * Nx*Ny*6, each Ex, Ey, Ez, Hx, Hy, Hz
 *           (the FieldAll components of data.jl:86-88), from eigen.jl:56-61,
 *           with (Hx, Hy) of unit 2-norm and its largest entry real positive.
 *   ncv <= 0: max(20, 2 nev + 1); at most 64.  tol, maxiter as for
 *   fdfd_eigenfrequency.  precision: FDFD_FP64 or FDFD_FP32_IR (the factors
 *   of the two diagonal blocks; FP64 FGMRES refines every application). */
int fdfd_vector_eigenmode(fdfd_handle* h, double omega, fdfd_complex neff,
                          int nev, int ncv, int precision, double tol,
                          int maxiter, fdfd_complex* beta, fdfd_complex* fields,
                          fdfd_stats* stats);

/* ---- modulated (MF-FDFD) driven solve ---------------------------------- */
/* solve(d::ModulatedDevice) of src/solver/modulation.jl:35-119 at one
 * carrier frequency omega (the reference loops over d.omega): TM only.  With
 * K = nsidebands and wn_j = omega + j*Omega (j = -K..K) the coupled system is
 *   A_j = Sxb dxb (1/mu0') Sxf dxf + Syb dyb (1/mu0') Syf dyf + wn_j^2 eps0' eps
 *   row block j += 0.5 wn_j^2 eps0' (conj(deps) x_{j+1} + deps x_{j-1}),
 *   b = i omega src in the centre block (j = 0),
 * with the stretch factors at omega (sharedpml != 0) or at wn_j.  The handle's
 * grid and eps (fdfd_set_grid, fdfd_set_eps) are used; deps and src are
 * Nx*Ny.  The 2K+1 sideband operators are factored (precision as in
 * fdfd_factor) and FP64 FGMRES on the coupled operator, preconditioned by
 * them, reaches ||b - A x|| <= tol ||b||.
 *   out: (2K+1) * Nx*Ny*3, sideband j = -K..K in order, each Ez, Hx, Hy with
 *        Hx = -Syf dyf Ez / (i wn_j mu0'), Hy = Sxf dxf Ez / (i wn_j mu0')
 *        (modulation.jl:108-109).
 * Every wn_j must be > 0; K <= 16.  stats.t_factor_ms is the batched
 * factorization, stats.t_solve_ms the whole call's device time. */
int fdfd_modulated_solve(fdfd_handle* h, double omega, double Omega,
                         int nsidebands, int sharedpml, const fdfd_complex* deps,
                         const fdfd_complex* src, int precision, double tol,
                         fdfd_complex* out, fdfd_stats* stats);

/* ---- Poynting vector and flux (post-processing of fields) --------------- */
/* poynting(field) (src/flux.jl:15-30) on a Nx*Ny*3 field array (Ez, Hx, Hy
 * for TM; Hz, Ex, Ey for TE; the fdfd_solve layout).  S: Sx then Sy, Nx*Ny
 * doubles each.
 *   TM: Sx = -1/2 Re(avg_x(Ez) conj(Hy)), Sy = 1/2 Re(avg_y(Ez) conj(Hx))
 *   TE: Sx = 1/2 Re(Ey conj(avg_x(Hz))), Sy = -1/2 Re(Ex conj(avg_y(Hz)))
 * avg = grid_average (src/grid.jl:157-162).  The reference's TE branch reads
 * an undefined field (flux.jl:26-27); these are its TE analogue.  The grid is
 * explicit: these calls never touch the handle's grid or factors. */
int fdfd_poynting(fdfd_handle* h, int pol, int Nx, int Ny,
                  const fdfd_complex* fields, double* S);
int fdfd_poynting_d(fdfd_handle* h, int pol, int Nx, int Ny,
                    const void* d_fields, void* d_S);

/* A flux plane: normal 1 (x) or 2 (y); index = the 0-based column (x normal)
 * or row (y normal), averaged with index + 1 as in flux.jl:42-44; cells lo..hi
 * (0-based, inclusive) along the plane. */
typedef struct {
    int normal;
    int index;
    int lo;
    int hi;
} fdfd_flux_plane;

/* flux_surface_integral(field, center, width, normal) (src/flux.jl:37-66)
 * for nplanes planes at once, one value each: the normal Poynting component
 * summed along the plane times the cell size along it.  x-normal TM is the
 * reference's formula; x-normal TE (broken in the reference, :49-54) and the
 * y-normal cases (unimplemented there, :58-65) are completed the same way. */
int fdfd_flux(fdfd_handle* h, int pol, int Nx, int Ny, double dx, double dy,
              const fdfd_complex* fields, int nplanes,
              const fdfd_flux_plane* planes, double* flux);
int fdfd_flux_d(fdfd_handle* h, int pol, int Nx, int Ny, double dx, double dy,
                const void* d_fields, int nplanes,
                const fdfd_flux_plane* planes, double* flux);
/* fdfd_flux_d without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * plane table and the per-plane sums go through a static per-GPU scratch of
 * 96 KiB instead of a per-call workspace, so nplanes <= 4096
 * (FDFD_ERR_INVALID above).  Same results as fdfd_flux_d.  Calls that overlap
 * on one GPU share that scratch and race. */
int fdfd_flux_d_noalloc(fdfd_handle* h, int pol, int Nx, int Ny, double dx,
                        double dy, const void* d_fields, int nplanes,
                        const fdfd_flux_plane* planes, double* flux);

/* ---- batched driven solves (many grids x frequencies x sources) --------- */
/* One job = one grid + eps + nomega frequencies + nsrc sources (each source
 * is solved at every frequency).  fdfd_batch_run factors all (job, omega)
 * pairs with the batched block kernels, in groups of frequencies (at most
 * FDFD_BATCH_MEMORY_GB, default 32, GB of factors per group), and solves the
 * sources of each frequency together; the next group is factored on a
 * second stream while the current one is solved.  Results are read back
 * with fdfd_batch_result.  Uneven grid sizes may be mixed.  The factor
 * memory of the run (two groups) stays allocated on the handle for later
 * batches until fdfd_destroy: freeing and reallocating tens of GB costs
 * seconds of slower GPU work while the driver clears the memory.
 * fdfd_batch_run works on the handle: like fdfd_set_grid, it releases every
 * factor the handle holds (their ids become FDFD_ERR_NOTFOUND) and leaves
 * the handle without a grid; call fdfd_set_grid / fdfd_set_eps again before
 * the next fdfd_factor. */
typedef struct fdfd_batch fdfd_batch;
int fdfd_batch_create(fdfd_handle* h, fdfd_batch** out);
int fdfd_batch_add(fdfd_batch* b, int Nx, int Ny, double dx, double dy,
                   int npml_x, int npml_y, double L0, int pol,
                   const fdfd_complex* eps, int nomega, const double* omegas,
                   int nsrc, const fdfd_complex* srcs, int* job_id);
int fdfd_batch_run(fdfd_batch* b, int precision, double tol, fdfd_stats* stats);
/* fields of (job, omega index, source index): Nx*Ny*3 complex. */
int fdfd_batch_result(fdfd_batch* b, int job_id, int iomega, int isrc,
                      fdfd_complex* fields);
int fdfd_batch_destroy(fdfd_batch* b);

/* ---- tuning and profiling ----------------------------------------------- */
/* Per-kernel device time accumulated with HIP events on the handle's stream
 * while profiling is enabled.  kernel: 0 = trailing GJE update, 1 = panel,
 * 2 = Schur formation, 3 = sweep GEMV, 4 = stencil/residual, 5 = Z build,
 * 6 = complex GEMM, 7 = operator assembly. */
int fdfd_profile_enable(fdfd_handle* h, int enable);
int fdfd_profile_read(fdfd_handle* h, int kernel, double* total_ms,
                      long long* launches, double* flops, double* bytes);
/* Panel width override (0 = default 32).  force_full_search != 0 makes every
 * panel search the whole column for its pivots (no candidate window). */
int fdfd_set_tuning(fdfd_handle* h, int panel_width, int force_full_search);

/* Autotuner (no counterpart in the reference).  Times a whole step (assemble
 * + factor + refined solve of a point source to 1e-10) on the handle's grid
 * and eps for every variant of the factorization kernels -- fp16x3 or FP32
 * products for complex64 GEMMs and trailing updates, one or two panels per
 * trailing pass, 8 or 4 waves and three- or four-multiplication products in
 * the FP32/FP64 GEMM -- and keeps the fastest on the handle.  With the
 * environment variable FDFD_TUNE_CACHE naming a file, the choice is recorded
 * there per (line length, line count, precision, pol), and handles read it
 * at factorization time (unless variant switches are set in the
 * environment).  stats.t_factor_ms: the winning step time.
 * fdfd_get_tuning: knobs[0..4] = gemm fp16x3, trailing fp16x3, fused panels,
 * GEMM waves, GEMM three-multiplication. */
int fdfd_autotune(fdfd_handle* h, int pol, double omega, int precision,
                  fdfd_stats* stats);
int fdfd_get_tuning(fdfd_handle* h, int* knobs, int nknobs);

/* Failure injection for testing the error paths (no counterpart in the
 * reference, whose only failure handling is the solver fallback of
 * solver.jl:19-28).  One-shot: the next operation of the named kind on the
 * handle fails as stated, then the handle behaves normally.
 *   FDFD_INJECT_PIVOT_FALLBACK  next factorization: the windowed pivot pass of
 *                               its first factor (of a batched build: the
 *                               first frequency) is treated as failing its
 *                               verification, so that factor is redone with
 *                               the full-column search (stats.pivot_fallback)
 *   FDFD_INJECT_SINGULAR        next factorization reports a zero pivot in
 *                               line block `arg` (FDFD_ERR_SINGULAR)
 *   FDFD_INJECT_NOCONV          next refined solve runs at most `arg` FGMRES
 *                               iterations and no restart (FDFD_ERR_NOCONV
 *                               unless that reaches the tolerance)
 *   FDFD_INJECT_OOM             next factorization fails its storage
 *                               allocation (FDFD_ERR_OOM) */
enum {
    FDFD_INJECT_NONE = 0,
    FDFD_INJECT_PIVOT_FALLBACK = 1,
    FDFD_INJECT_SINGULAR = 2,
    FDFD_INJECT_NOCONV = 3,
    FDFD_INJECT_OOM = 4
};
int fdfd_debug_inject(fdfd_handle* h, int what, int arg);

#ifdef __cplusplus
}
#endif
#endif /* FDFD_H */
