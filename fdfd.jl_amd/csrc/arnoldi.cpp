// Thick-restart Arnoldi on a device operator OP (vectors of N complex
// doubles), the eigensolver behind fdfd_eigenfrequency and
// fdfd_vector_eigenmode.  It plays the part of Arpack.eigs with a shift
// (src/solver/eigen.jl:26,49,86,104): OP is the shift-inverted operator and
// the wanted Ritz values are the largest in magnitude (or real part).
//
// Arnoldi vectors live on the device (FP64, CGS2 orthogonalisation); the
// small projected matrix is solved on the host (small_eig.h).  Restarts keep
// the span of the `keep` wanted Ritz vectors (thick restart).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "runtime.h"

namespace fdfd {

int rt_norm(Handle* h, const double2* x, long N, double* out);
int rt_multidot(Handle* h, const double2* V, long N, int k, const double2* w, cd* out);
int rt_update(Handle* h, double2* w, const double2* V, long N, int k, const cd* c, double sign);
int small_eig(int n, std::vector<cd> A, std::vector<cd>& evals, std::vector<cd>& evecs);

int rt_arnoldi(Handle* h, long N, int nev, int ncv, int which, double tol, int maxiter,
               const std::function<int(const double2*, double2*)>& op, double2* ritz,
               std::vector<cd>& theta_out, int* nconv_out, int* restarts_out,
               const double* op_error) {
    double2* Vb = nullptr;  // ncv + 1 Arnoldi vectors
    double2* Wb = nullptr;  // restart workspace
    auto cleanup = [&]() {
        if (Vb) (void)hipFree(Vb);
        if (Wb) (void)hipFree(Wb);
    };
    if (hipMalloc(&Vb, sizeof(double2) * N * (ncv + 1)) != hipSuccess ||
        hipMalloc(&Wb, sizeof(double2) * N * ncv) != hipSuccess) {
        cleanup();
        return fail(h, FDFD_ERR_OOM, "Arnoldi basis allocation failed");
    }
    int rc;
    auto vec = [&](int i) { return Vb + (long)i * N; };
    // deterministic start vector
    {
        std::vector<double2> v0(N);
        uint64_t s = 0x9e3779b97f4a7c15ULL;
        for (long i = 0; i < N; ++i) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            double a = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            double b = (double)(s >> 11) / 9007199254740992.0 - 0.5;
            v0[i] = make_double2(a, b);
        }
        if ((rc = stream_copy(h, vec(0), v0.data(), sizeof(double2) * N, hipMemcpyHostToDevice))) {
            cleanup();
            return rc;
        }
        double nv;
        if ((rc = rt_norm(h, vec(0), N, &nv))) { cleanup(); return rc; }
        launch_scale_copy(h->stream, vec(0), vec(0), make_double2(1.0 / nv, 0), N);
    }
    const int m = ncv;
    std::vector<cd> H((size_t)(m + 1) * m, 0.0);  // row-major (m+1) x m
    auto Hat = [&](int i, int j) -> cd& { return H[(size_t)i * m + j]; };
    std::vector<cd> hc(m + 1), hc2(m + 1);
    int start = 0, nconv = 0, restarts = 0;
    std::vector<cd> theta, Y;
    std::vector<int> order;
    const int keep = std::min(m - 1, nev + std::max(2, (m - nev) / 2));
    for (;;) {
        for (int j = start; j < m; ++j) {
            double2* w = vec(j + 1);
            if ((rc = op(vec(j), w))) { cleanup(); return rc; }
            if ((rc = rt_multidot(h, Vb, N, j + 1, w, hc.data()))) { cleanup(); return rc; }
            if ((rc = rt_update(h, w, Vb, N, j + 1, hc.data(), -1.0))) { cleanup(); return rc; }
            if ((rc = rt_multidot(h, Vb, N, j + 1, w, hc2.data()))) { cleanup(); return rc; }
            if ((rc = rt_update(h, w, Vb, N, j + 1, hc2.data(), -1.0))) { cleanup(); return rc; }
            for (int i = 0; i <= j; ++i) Hat(i, j) = hc[i] + hc2[i];
            double hn;
            if ((rc = rt_norm(h, w, N, &hn))) { cleanup(); return rc; }
            Hat(j + 1, j) = hn;
            if (hn > 0) launch_scale_copy(h->stream, w, w, make_double2(1.0 / hn, 0), N);
        }
        std::vector<cd> Hm((size_t)m * m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) Hm[(size_t)i * m + j] = Hat(i, j);
        if (small_eig(m, Hm, theta, Y) != 0) {
            cleanup();
            return fail(h, FDFD_ERR_NOCONV, "projected eigenproblem did not converge");
        }
        order.resize(m);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (which == FDFD_WHICH_LR) return theta[a].real() > theta[b].real();
            return std::abs(theta[a]) > std::abs(theta[b]);
        });
        // residual of Ritz pair i: |H[m, :] . y_i|
        // a Ritz residual cannot go below the relative error of the OP
        // applications (op_error, when the caller tracks it)
        const double tol_eff = op_error ? std::max(tol, 10.0 * *op_error) : tol;
        nconv = 0;
        for (int q = 0; q < nev; ++q) {
            int i = order[q];
            cd r = 0;
            for (int j = 0; j < m; ++j) r += Hat(m, j) * Y[(size_t)j * m + i];
            if (std::abs(r) <= tol_eff * std::abs(theta[i])) ++nconv;
        }
        if (nconv >= nev || restarts >= maxiter) break;
        ++restarts;
        // thick restart: orthonormal basis of the `keep` wanted Ritz vectors
        std::vector<cd> Qk((size_t)m * keep);
        for (int q = 0; q < keep; ++q)
            for (int i = 0; i < m; ++i) Qk[(size_t)i * keep + q] = Y[(size_t)i * m + order[q]];
        for (int q = 0; q < keep; ++q) {
            for (int pass = 0; pass < 2; ++pass)
                for (int p = 0; p < q; ++p) {
                    cd s = 0;
                    for (int i = 0; i < m; ++i) s += std::conj(Qk[(size_t)i * keep + p]) * Qk[(size_t)i * keep + q];
                    for (int i = 0; i < m; ++i) Qk[(size_t)i * keep + q] -= s * Qk[(size_t)i * keep + p];
                }
            double nq = 0;
            for (int i = 0; i < m; ++i) nq += std::norm(Qk[(size_t)i * keep + q]);
            nq = std::sqrt(nq);
            for (int i = 0; i < m; ++i) Qk[(size_t)i * keep + q] /= nq;
        }
        // V[:, 0:keep] = V_m Qk ; v_keep = v_m
        for (int q = 0; q < keep; ++q) {
            std::vector<cd> col(m);
            for (int i = 0; i < m; ++i) col[i] = Qk[(size_t)i * keep + q];
            launch_zero(h->stream, Wb + (long)q * N, N);
            if ((rc = rt_update(h, Wb + (long)q * N, Vb, N, m, col.data(), 1.0))) { cleanup(); return rc; }
        }
        if (hipMemcpyAsync(vec(keep), vec(m), sizeof(double2) * N, hipMemcpyDeviceToDevice, h->stream) !=
                hipSuccess ||
            hipMemcpyAsync(Vb, Wb, sizeof(double2) * N * keep, hipMemcpyDeviceToDevice, h->stream) !=
                hipSuccess) {
            cleanup();
            return fail(h, FDFD_ERR_HIP, "restart copy failed");
        }
        // H_new = Qk^H H_m Qk, row keep = H[m, :] Qk
        std::vector<cd> HQ((size_t)m * keep, 0.0), Hn((size_t)keep * keep, 0.0), hr(keep, 0.0);
        for (int i = 0; i < m; ++i)
            for (int q = 0; q < keep; ++q)
                for (int j = 0; j < m; ++j) HQ[(size_t)i * keep + q] += Hat(i, j) * Qk[(size_t)j * keep + q];
        for (int p = 0; p < keep; ++p)
            for (int q = 0; q < keep; ++q)
                for (int i = 0; i < m; ++i)
                    Hn[(size_t)p * keep + q] += std::conj(Qk[(size_t)i * keep + p]) * HQ[(size_t)i * keep + q];
        for (int q = 0; q < keep; ++q)
            for (int j = 0; j < m; ++j) hr[q] += Hat(m, j) * Qk[(size_t)j * keep + q];
        std::fill(H.begin(), H.end(), cd(0));
        for (int p = 0; p < keep; ++p)
            for (int q = 0; q < keep; ++q) Hat(p, q) = Hn[(size_t)p * keep + q];
        for (int q = 0; q < keep; ++q) Hat(keep, q) = hr[q];
        start = keep;
    }
    theta_out.resize(nev);
    for (int q = 0; q < nev; ++q) theta_out[q] = theta[order[q]];
    // Ritz vectors of the nev wanted values: unit 2-norm, largest entry real > 0
    if (ritz) {
        std::vector<double2> hx(N);
        for (int q = 0; q < nev; ++q) {
            std::vector<cd> col(m);
            for (int i = 0; i < m; ++i) col[i] = Y[(size_t)i * m + order[q]];
            double2* x = ritz + (long)q * N;
            launch_zero(h->stream, x, N);
            if ((rc = rt_update(h, x, Vb, N, m, col.data(), 1.0))) { cleanup(); return rc; }
            if ((rc = stream_copy(h, hx.data(), x, sizeof(double2) * N, hipMemcpyDeviceToHost))) {
                cleanup();
                return rc;
            }
            double nx = 0, best = -1;
            cd ph = 1;
            for (long i = 0; i < N; ++i) {
                double a2 = hx[i].x * hx[i].x + hx[i].y * hx[i].y;
                nx += a2;
                if (a2 > best) { best = a2; ph = cd(hx[i].x, hx[i].y); }
            }
            cd scale = std::conj(ph / std::abs(ph)) / std::sqrt(nx);
            launch_scale_copy(h->stream, x, x, make_double2(scale.real(), scale.imag()), N);
        }
    }
    *nconv_out = nconv;
    *restarts_out = restarts;
    cleanup();
    return FDFD_OK;
}

}  // namespace fdfd
