// fdfd_eigenmode: batched 1-D eigenmodes of eps slices, the mode-source
// solver behind get_modes / setup_mode! (src/device.jl:118-149) and
// eigenmode(d::Device{1}, pol, neff, nev) (src/solver/eigen.jl:6-29).
// One workgroup per slice (k_mode.hip); the host packs the slices, sets the
// shift sigma = (omega/c0' neff)^2 and the 1-D grid's dx = (n dh - 0)/n
// (Grid(n, [0 n*dh]), src/grid.jl:53-58,68-70), and reads the results back.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.h"
#include "tri_qr.h"

using namespace fdfd;

namespace {
constexpr int MODE_MAX_N = 16384;  // longest slice (one lane runs the O(n) solves)
}

extern "C" int fdfd_eigenmode(fdfd_handle* hh, int nprob, const fdfd_mode_problem* probs,
                              int nev, int ncv, double L0, double tol, int maxiter,
                              fdfd_complex* beta, fdfd_complex* vectors, int* nconv,
                              fdfd_stats* stats) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h) return FDFD_ERR_INVALID;
    if (nprob < 1 || !probs || !beta) return fail(h, FDFD_ERR_INVALID, "nprob >= 1, probs and beta required");
    if (nev < 1) return fail(h, FDFD_ERR_INVALID, "nev must be >= 1");
    if (ncv <= 0) ncv = std::max(20, 2 * nev + 1);  // Arpack.jl default
    if (ncv > MODE_NCV)
        return fail(h, FDFD_ERR_INVALID, "ncv must be <= " + std::to_string(MODE_NCV));
    if (!(L0 > 0)) L0 = 1e-6;
    if (!(tol > 0)) tol = 1e-12;
    if (maxiter <= 0) maxiter = 300;
    const double eps0n = EPS0 * L0, mu0n = MU0 * L0;
    const double c0n = std::sqrt(1.0 / EPS0 / MU0) / L0;
    std::vector<size_t> off(nprob + 1, 0);
    std::vector<int> mp(nprob);
    for (int p = 0; p < nprob; ++p) {
        const fdfd_mode_problem& P = probs[p];
        const std::string tag = "mode problem " + std::to_string(p) + ": ";
        if (P.n < 3 || P.n > MODE_MAX_N)
            return fail(h, FDFD_ERR_INVALID, tag + "n must be in [3, " + std::to_string(MODE_MAX_N) + "]");
        if (P.pol != FDFD_TM && P.pol != FDFD_TE) return fail(h, FDFD_ERR_INVALID, tag + "pol must be TM or TE");
        if (!(P.omega > 0) || !(P.dh > 0)) return fail(h, FDFD_ERR_INVALID, tag + "omega and dh must be > 0");
        if (!P.eps) return fail(h, FDFD_ERR_INVALID, tag + "eps is NULL");
        mp[p] = std::min(ncv, P.n);
        if (nev >= mp[p]) return fail(h, FDFD_ERR_INVALID, tag + "nev must be < min(ncv, n)");
        off[p + 1] = off[p] + P.n;
    }
    (void)hipSetDevice(h->device);
    const size_t ntot = off[nprob];
    // one device buffer: eps | V | W | vec | beta (double2), tri work (double), info (int)
    size_t nV = 0, nW = 0, nT = 0;
    for (int p = 0; p < nprob; ++p) {
        nV += (size_t)(mp[p] + 1) * probs[p].n;
        nW += (size_t)mp[p] * probs[p].n;
        nT += (size_t)tri::work_doubles(probs[p].n);
    }
    const size_t n2 = ntot + nV + nW + ntot * nev + (size_t)nprob * nev;
    const size_t bytes = n2 * sizeof(double2) + nT * sizeof(double) + (size_t)nprob * 4 * sizeof(int) +
                         (size_t)nprob * sizeof(ModeJob) + alignof(ModeJob);
    char* buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) return fail(h, FDFD_ERR_OOM, "mode workspace allocation failed");
    struct Free {
        char* p;
        ~Free() { (void)hipFree(p); }
    } guard{buf};
    double2* d_eps = (double2*)buf;
    double2* d_V = d_eps + ntot;
    double2* d_W = d_V + nV;
    double2* d_vec = d_W + nW;
    double2* d_beta = d_vec + ntot * nev;
    double* d_trw = (double*)(d_beta + (size_t)nprob * nev);
    int* d_info = (int*)(d_trw + nT);
    ModeJob* d_jobs = (ModeJob*)(d_info + (size_t)nprob * 4);
    // ModeJob holds doubles: keep the table 8-byte aligned
    if (((uintptr_t)d_jobs) % alignof(ModeJob) != 0)
        d_jobs = (ModeJob*)(((uintptr_t)d_jobs + alignof(ModeJob) - 1) & ~(uintptr_t)(alignof(ModeJob) - 1));
    std::vector<double2> heps(ntot);
    std::vector<ModeJob> jobs(nprob);
    size_t oV = 0, oW = 0, oT = 0;
    for (int p = 0; p < nprob; ++p) {
        const fdfd_mode_problem& P = probs[p];
        for (int i = 0; i < P.n; ++i) heps[off[p] + i] = make_double2(P.eps[i].re, P.eps[i].im);
        const double dx = ((double)P.n * P.dh - 0.0) / P.n;
        const std::complex<double> kb = P.omega / c0n * std::complex<double>(P.neff.re, P.neff.im);
        const std::complex<double> sg = kb * kb;
        ModeJob& J = jobs[p];
        J.eps = d_eps + off[p];
        J.V = d_V + oV;
        J.W = d_W + oW;
        J.trw = d_trw + oT;
        J.vec = d_vec + off[p] * nev;
        J.beta = d_beta + (size_t)p * nev;
        J.info = d_info + 4 * p;
        J.sigma = make_double2(sg.real(), sg.imag());
        J.k0sq = P.omega * P.omega * mu0n * eps0n;
        J.idx2 = 1.0 / (dx * dx);
        J.n = P.n;
        J.pol = P.pol;
        J.ncv = mp[p];
        oV += (size_t)(mp[p] + 1) * P.n;
        oW += (size_t)mp[p] * P.n;
        oT += (size_t)tri::work_doubles(P.n);
    }
    COPY(h, d_eps, heps.data(), ntot * sizeof(double2), hipMemcpyHostToDevice);
    COPY(h, d_jobs, jobs.data(), nprob * sizeof(ModeJob), hipMemcpyHostToDevice);
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    (void)hipEventRecord(e0, h->stream);
    launch_mode(h->stream, d_jobs, nprob, nev, tol, maxiter);
    (void)hipEventRecord(e1, h->stream);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return fail(h, FDFD_ERR_HIP, std::string("k_mode launch: ") + hipGetErrorString(le));
    }
    std::vector<int> info((size_t)nprob * 4);
    std::vector<double2> hb((size_t)nprob * nev);
    int rc = stream_copy(h, info.data(), d_info, info.size() * sizeof(int), hipMemcpyDeviceToHost);
    if (!rc) rc = stream_copy(h, hb.data(), d_beta, hb.size() * sizeof(double2), hipMemcpyDeviceToHost);
    if (!rc && vectors)
        rc = stream_copy(h, vectors, d_vec, ntot * nev * sizeof(double2), hipMemcpyDeviceToHost);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    for (size_t i = 0; i < hb.size(); ++i) {
        beta[i].re = hb[i].x;
        beta[i].im = hb[i].y;
    }
    int worst = nev, restarts = 0, applies = 0, bad = -1, status = MODE_OK;
    for (int p = 0; p < nprob; ++p) {
        const int* I = &info[(size_t)4 * p];
        if (nconv) nconv[p] = I[1];
        worst = std::min(worst, I[1]);
        restarts = std::max(restarts, I[2]);
        applies += I[3];
        if (bad < 0 && (I[0] != MODE_OK || I[1] < nev)) {
            bad = p;
            status = I[0];
        }
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = applies;
        stats->restarts = restarts;
        stats->t_solve_ms = ms;
        stats->singular_line = (status == MODE_SINGULAR) ? bad : -1;
        stats->nconv = worst;
    }
    if (bad >= 0) {
        const std::string tag = "mode problem " + std::to_string(bad) + ": ";
        if (status == MODE_SINGULAR)
            return fail(h, FDFD_ERR_SINGULAR, tag + "the shift is an eigenvalue (A - sigma I is singular)");
        if (status == MODE_EIG_FAIL)
            return fail(h, FDFD_ERR_NOCONV, tag + "projected eigenproblem did not converge");
        return fail(h, FDFD_ERR_NOCONV, tag + "Arnoldi: " + std::to_string(info[(size_t)4 * bad + 1]) +
                                            " of " + std::to_string(nev) + " eigenpairs converged");
    }
    return FDFD_OK;
}
