// Host entry of the small dense complex eigensolver (small_eig.h) used by the
// eigenfrequency Arnoldi driver: std::vector storage around the shared
// host/device implementation.
#include <complex>
#include <vector>

#include "small_eig.h"

namespace fdfd {
using cd = std::complex<double>;

// A row-major n x n.  evals[k], evecs[i * n + k] (column k = eigenvector k,
// unit 2-norm).  Returns 0, or -1 if the QR iteration did not converge.
int small_eig(int n, std::vector<cd> A, std::vector<cd>& evals, std::vector<cd>& evecs) {
    std::vector<eig::cz> a((size_t)n * n), q((size_t)n * n), ev(n), V((size_t)n * n), z(n), v(n);
    for (size_t i = 0; i < a.size(); ++i) a[i] = eig::mkz(A[i].real(), A[i].imag());
    const int rc = eig::small_eig(n, a.data(), q.data(), ev.data(), V.data(), z.data(), v.data());
    if (rc) return rc;
    evals.resize(n);
    evecs.resize((size_t)n * n);
    for (int i = 0; i < n; ++i) evals[i] = cd(ev[i].re, ev[i].im);
    for (size_t i = 0; i < V.size(); ++i) evecs[i] = cd(V[i].re, V[i].im);
    return 0;
}

}  // namespace fdfd
