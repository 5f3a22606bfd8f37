// CPU execution of the block cyclic reduction plan (fdfd.jl_amd/csrc/bcr_plan.h)
// with dense complex<double> blocks.  Each step follows the job semantics the
// device kernels implement (FormJob, GemmJob, MvJob, DiagJob in fdfd_common.h,
// Gauss-Jordan output X[a, b] = Y[prow[a], iprow[b]] with a random pivot
// permutation), on random periodic 5-point operators of every small shape.
// The plan is a direct solver, so the solve of A x = b must leave a residual
// at round-off level.  Prints one line per case; exit status 1 on failure.
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "bcr_plan.h"

using cd = std::complex<double>;
using namespace fdfd::bcr;

struct Op {
    int n, m;
    // planes indexed [line * n + i]: 0 W (i-1), 1 E (i+1), 2 S (line-1), 3 N (line+1), 4 C
    std::vector<cd> pl[5];
};

static std::vector<cd> apply_op(const Op& A, const std::vector<cd>& x) {
    const int n = A.n, m = A.m;
    std::vector<cd> y(x.size());
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < n; ++i) {
            const int k = j * n + i;
            const int ip = (i + 1) % n, im = (i + n - 1) % n;
            const int jp = (j + 1) % m, jm = (j + m - 1) % m;
            y[k] = A.pl[4][k] * x[k] + A.pl[0][k] * x[j * n + im] + A.pl[1][k] * x[j * n + ip] +
                   A.pl[2][k] * x[jm * n + i] + A.pl[3][k] * x[jp * n + i];
        }
    return y;
}

// dense inverse by Gauss-Jordan with partial pivoting
static bool invert(std::vector<cd>& a, int n) {
    std::vector<cd> inv(n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(a[r * n + c]) > std::abs(a[p * n + c])) p = r;
        if (std::abs(a[p * n + c]) == 0.0) return false;
        for (int k = 0; k < n; ++k) {
            std::swap(a[c * n + k], a[p * n + k]);
            std::swap(inv[c * n + k], inv[p * n + k]);
        }
        const cd d = 1.0 / a[c * n + c];
        for (int k = 0; k < n; ++k) {
            a[c * n + k] *= d;
            inv[c * n + k] *= d;
        }
        for (int r = 0; r < n; ++r)
            if (r != c) {
                const cd f = a[r * n + c];
                for (int k = 0; k < n; ++k) {
                    a[r * n + k] -= f * a[c * n + k];
                    inv[r * n + k] -= f * inv[c * n + k];
                }
            }
    }
    a = inv;
    return true;
}

static double run_case(int n, int m, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    Op A{n, m, {}};
    for (int q = 0; q < 5; ++q) {
        A.pl[q].resize((size_t)n * m);
        for (auto& v : A.pl[q]) v = cd(U(rng), U(rng));
    }
    for (auto& v : A.pl[4]) v += cd(6.0, 1.0);  // keep every block well conditioned
    const Plan P = make_plan(m);
    std::vector<std::vector<cd>> blk(P.nblocks, std::vector<cd>((size_t)n * n, cd(0)));
    std::vector<std::vector<int>> prow(m, std::vector<int>(n)), iprow(m, std::vector<int>(n));
    auto dvec = [&](const SDiag& d, int i) {
        cd v = 0;
        for (int q = 0; q < 2; ++q)
            if (d.line >= 0 && d.pl[q] >= 0) v += A.pl[d.pl[q] ? 3 : 2][d.line * n + i];
        return v;
    };
    auto X = [&](int j, int a, int b) { return blk[j][(size_t)prow[j][a] * n + iprow[j][b]]; };
    for (const Step& st : P.fsteps) {
        for (int c = 0; c < st.count; ++c) {
            if (st.kind == ST_TRI) {
                const STri& t = P.tris[st.first + c];
                std::vector<cd> a((size_t)n * n, cd(0));
                for (int i = 0; i < n; ++i) {
                    const int k = t.line * n + i;
                    a[(size_t)i * n + i] += A.pl[4][k] + dvec(t.dd, i);
                    a[(size_t)i * n + (i + 1) % n] += A.pl[1][k];
                    a[(size_t)i * n + (i + n - 1) % n] += A.pl[0][k];
                }
                if (!invert(a, n)) return 1e300;
                blk[t.line] = a;
                for (int i = 0; i < n; ++i) prow[t.line][i] = iprow[t.line][i] = i;
            } else if (st.kind == ST_FORM) {
                const SForm& f = P.forms[st.first + c];
                std::vector<cd> out((size_t)n * n, cd(0));
                for (int a = 0; a < n; ++a)
                    for (int b = 0; b < n; ++b) {
                        cd v = 0;
                        if (f.own >= 0) {
                            const int k = f.own * n + a;
                            if (b == a) v += A.pl[4][k];
                            if (b == (a + 1) % n) v += A.pl[1][k];
                            if (b == (a + n - 1) % n) v += A.pl[0][k];
                        }
                        if (b == a) v += dvec(f.dd, a);
                        for (int t = 0; t < f.nt; ++t)
                            v -= dvec(f.rs[t], a) * X(f.jl[t], a, b) * dvec(f.cs[t], b);
                        out[(size_t)a * n + b] = v;
                    }
                blk[f.out] = out;
            } else if (st.kind == ST_GJ) {
                const int j = P.gj[st.first + c];
                std::vector<cd> x = blk[j];
                if (!invert(x, n)) return 1e300;
                std::vector<int> p(n);
                for (int i = 0; i < n; ++i) p[i] = i;
                std::shuffle(p.begin(), p.end(), rng);
                for (int i = 0; i < n; ++i) {
                    prow[j][i] = p[i];
                    iprow[j][p[i]] = i;
                }
                for (int a = 0; a < n; ++a)
                    for (int b = 0; b < n; ++b) blk[j][(size_t)prow[j][a] * n + iprow[j][b]] = x[(size_t)a * n + b];
            } else {
                const SGemm& g = P.gemms[st.first + c];
                std::vector<cd> out = g.acc ? blk[g.out] : std::vector<cd>((size_t)n * n, cd(0));
                for (int t = 0; t < g.nt; ++t) {
                    const auto& Am = blk[g.A[t]];
                    const auto& Bm = blk[g.B[t]];
                    for (int i = 0; i < n; ++i)
                        for (int k = 0; k < n; ++k) {
                            const int br = g.bmap[t] >= 0 ? prow[g.bmap[t]][k] : k;
                            const cd aik = Am[(size_t)i * n + k];
                            const cd f = g.sub ? -aik : aik;
                            for (int jj = 0; jj < n; ++jj) out[(size_t)i * n + jj] += f * Bm[(size_t)br * n + jj];
                        }
                }
                blk[g.out] = out;
            }
        }
    }
    // solve A x = b through the program
    std::vector<cd> b((size_t)n * m);
    for (auto& v : b) v = cd(U(rng), U(rng));
    std::vector<cd> vx = b, vt((size_t)n * m, cd(0));
    auto V = [&](int v) -> std::vector<cd>& { return v == VX ? vx : vt; };
    for (const Step& st : P.ssteps) {
        for (int c = 0; c < st.count; ++c) {
            if (st.kind == ST_MV) {
                const SMv& s = P.mvs[st.first + c];
                std::vector<cd> z(n, cd(0));
                for (int t = 0; t < s.nt; ++t) {
                    const auto& Mm = blk[s.M[t]];
                    const auto& xv = V(s.xv[t]);
                    for (int r = 0; r < n; ++r)
                        for (int cc = 0; cc < n; ++cc) {
                            const int src = s.imap[t] >= 0 ? prow[s.imap[t]][cc] : cc;
                            z[r] += Mm[(size_t)r * n + cc] * xv[s.xl[t] * n + src];
                        }
                }
                auto& out = V(s.ov);
                std::vector<cd> base = s.bl >= 0 ? V(s.bv) : std::vector<cd>();
                for (int r = 0; r < n; ++r) {
                    const int o = s.omap >= 0 ? iprow[s.omap][r] : r;
                    out[s.ol * n + o] = s.bl >= 0 ? base[s.bl * n + o] - z[r] : z[r];
                }
            } else if (st.kind == ST_TRS) {
                // line solve by the first-level inverse (the device runs the
                // stored Givens QR; here the dense inverse of the same block)
                const STrs& s = P.trss[st.first + c];
                if (P.tris[s.tri].line != s.line) return 1e300;
                const auto& Mm = blk[s.line];
                const auto& xv = V(s.xv);
                std::vector<cd> z(n, cd(0));
                for (int r = 0; r < n; ++r)
                    for (int cc = 0; cc < n; ++cc) z[r] += Mm[(size_t)r * n + cc] * xv[s.line * n + cc];
                auto& out = V(s.ov);
                for (int r = 0; r < n; ++r) out[s.line * n + r] = z[r];
            } else {
                const SDg& s = P.dgs[st.first + c];
                std::vector<cd> res(n);
                for (int i = 0; i < n; ++i) {
                    cd v = V(s.x0v)[s.x0l * n + i];
                    for (int t = 0; t < s.nt; ++t) v -= dvec(s.d[t], i) * V(s.xv[t])[s.xl[t] * n + i];
                    res[i] = v;
                }
                for (int i = 0; i < n; ++i) V(s.ov)[s.ol * n + i] = res[i];
            }
        }
    }
    const std::vector<cd> y = apply_op(A, vx);
    double rn = 0, bn = 0;
    for (size_t k = 0; k < b.size(); ++k) {
        rn += std::norm(y[k] - b[k]);
        bn += std::norm(b[k]);
    }
    return std::sqrt(rn / bn);
}

int main() {
    int bad = 0;
    const int ns[] = {1, 2, 3, 5, 8};
    const int ms[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 17, 23, 32, 33};
    unsigned seed = 1;
    for (int n : ns)
        for (int m : ms) {
            const double r = run_case(n, m, seed++);
            const bool ok = r < 1e-12;
            bad += !ok;
            std::printf("n=%d m=%d residual=%.3e %s\n", n, m, r, ok ? "ok" : "FAIL");
        }
    return bad ? 1 : 0;
}
