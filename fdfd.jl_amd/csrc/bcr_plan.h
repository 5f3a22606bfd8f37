// Symbolic plan of the block cyclic reduction (host-only C++, no HIP).
//
// The plan depends on the number of lines m only.  Block ids 0..m-1 are the
// lines' own blocks (line j is inverted in place in block j); coupling and
// product blocks are drawn from a pool and recycled once consumed.
// Factorization steps (in order):
//   ST_TRI   first-level inversion of the listed lines' own periodic
//            tridiagonal blocks (Givens QR; natural order, identity maps)
//   ST_FORM  first-level block formation from the coefficient planes
//   ST_GJ    batched Gauss-Jordan inversion of the listed lines
//   ST_GEMM  batched complex products out = [out] -/+ sum_t A_t B_t[prow_t]
// Solve steps act on two field-shaped vectors, VX (in/out) and VT (temp):
//   ST_MV    dense matrix-vector steps, ST_DIAG first-level diagonal steps,
//   ST_TRS   first-level line solves x_out = T_j^-1 x_in by the line's stored
//            Givens QR (T_j its periodic tridiagonal block, X_j = T_j^-1).
// With X_j = D_j^-1 stored as X[a, b] = Y_j[prow_j[a], iprow_j[b]]:
//   Z(u, j) = C(u, j) Y_j[prow_j, :]     (= C(u, j) X_j, columns by iprow_j)
//   D_s    -= Z(s, j) C(j, s)[prow_j, :] (= C(s, j) X_j C(j, s))
#pragma once

#include <algorithm>
#include <map>
#include <numeric>
#include <utility>
#include <vector>

namespace fdfd {
namespace bcr {

// sum of at most two coefficient planes of one line (plane 0: coupling to the
// previous line, plane 1: to the next line)
struct SDiag {
    int line = -1;
    int pl[2] = {-1, -1};
};
struct SCoup {
    bool dense = false;
    SDiag d;
    int blk = -1;
};
struct STri {
    int line = -1;
    SDiag dd;  // added to the diagonal (only for a single line: its own wrap)
};
struct SForm {
    int out = -1, own = -1;
    SDiag dd;
    int nt = 0;
    int jl[2] = {-1, -1};
    SDiag rs[2], cs[2];
};
struct SGemm {
    int out = -1, acc = 0, sub = 1, nt = 0;
    int A[2] = {-1, -1}, B[2] = {-1, -1}, bmap[2] = {-1, -1};
};
enum Vec { VX = 0, VT = 1 };
struct SMv {
    int nt = 0;
    int M[2] = {-1, -1};
    int imap[2] = {-1, -1};  // line whose prow gathers the input
    int xv[2] = {0, 0}, xl[2] = {-1, -1};
    int bv = -1, bl = -1;
    int ov = 0, ol = -1;
    int omap = -1;  // line whose iprow scatters the output
};
struct STrs {
    int tri = -1, line = -1;  // index into tris, line
    int xv = 0, ov = 0;       // vectors read and written (same line)
};
struct SDg {
    int ov = 0, ol = -1, x0v = 0, x0l = -1;
    int nt = 0;
    int xv[2] = {0, 0}, xl[2] = {-1, -1};
    SDiag d[2];
};
enum StepKind { ST_FORM = 0, ST_GJ = 1, ST_GEMM = 2, ST_MV = 3, ST_DIAG = 4, ST_TRI = 5, ST_TRS = 6 };
struct Step {
    int kind, first, count;
};

struct Plan {
    int m = 0, nblocks = 0, max_gj = 0;
    std::vector<STri> tris;
    std::vector<SForm> forms;
    std::vector<int> gj;
    std::vector<SGemm> gemms;
    std::vector<Step> fsteps;
    std::vector<SMv> mvs;
    std::vector<SDg> dgs;
    std::vector<STrs> trss;
    std::vector<Step> ssteps;
};

inline Plan make_plan(int m) {
    Plan P;
    P.m = m;
    int next_blk = m;
    std::vector<int> freel;
    auto alloc = [&]() {
        if (!freel.empty()) {
            const int b = freel.back();
            freel.pop_back();
            return b;
        }
        return next_blk++;
    };
    auto step_form = [&](const std::vector<SForm>& v) {
        if (v.empty()) return;
        P.fsteps.push_back({ST_FORM, (int)P.forms.size(), (int)v.size()});
        P.forms.insert(P.forms.end(), v.begin(), v.end());
    };
    auto step_tri = [&](const std::vector<STri>& v) {
        P.fsteps.push_back({ST_TRI, (int)P.tris.size(), (int)v.size()});
        P.tris.insert(P.tris.end(), v.begin(), v.end());
        P.max_gj = std::max(P.max_gj, (int)v.size());
    };
    auto step_gj = [&](const std::vector<int>& v) {
        P.fsteps.push_back({ST_GJ, (int)P.gj.size(), (int)v.size()});
        P.gj.insert(P.gj.end(), v.begin(), v.end());
        P.max_gj = std::max(P.max_gj, (int)v.size());
    };
    auto step_gemm = [&](const std::vector<SGemm>& v) {
        if (v.empty()) return;
        P.fsteps.push_back({ST_GEMM, (int)P.gemms.size(), (int)v.size()});
        P.gemms.insert(P.gemms.end(), v.begin(), v.end());
    };
    // solve steps: forward ones are appended directly, backward ones are
    // collected per level and appended in reverse level order at the end
    struct SolveBlock {
        std::vector<SMv> mv;
        std::vector<SDg> dg;
        std::vector<STrs> ts;
        std::vector<Step> st;
    };
    auto add_mv = [](SolveBlock& b, const std::vector<SMv>& v) {
        if (v.empty()) return;
        b.st.push_back({ST_MV, (int)b.mv.size(), (int)v.size()});
        b.mv.insert(b.mv.end(), v.begin(), v.end());
    };
    auto add_dg = [](SolveBlock& b, const std::vector<SDg>& v) {
        if (v.empty()) return;
        b.st.push_back({ST_DIAG, (int)b.dg.size(), (int)v.size()});
        b.dg.insert(b.dg.end(), v.begin(), v.end());
    };
    auto add_ts = [](SolveBlock& b, const std::vector<STrs>& v) {
        if (v.empty()) return;
        b.st.push_back({ST_TRS, (int)b.ts.size(), (int)v.size()});
        b.ts.insert(b.ts.end(), v.begin(), v.end());
    };
    auto emit = [&](const SolveBlock& b) {
        for (const Step& s : b.st) {
            if (s.kind == ST_MV) {
                P.ssteps.push_back({ST_MV, (int)P.mvs.size(), s.count});
                P.mvs.insert(P.mvs.end(), b.mv.begin() + s.first, b.mv.begin() + s.first + s.count);
            } else if (s.kind == ST_TRS) {
                P.ssteps.push_back({ST_TRS, (int)P.trss.size(), s.count});
                P.trss.insert(P.trss.end(), b.ts.begin() + s.first, b.ts.begin() + s.first + s.count);
            } else {
                P.ssteps.push_back({ST_DIAG, (int)P.dgs.size(), s.count});
                P.dgs.insert(P.dgs.end(), b.dg.begin() + s.first, b.dg.begin() + s.first + s.count);
            }
        }
    };
    // x_j <- X_j x_j through the temporary vector
    auto apply_x = [](int j) {
        SMv v;
        v.nt = 1;
        v.M[0] = j;
        v.imap[0] = j;
        v.xv[0] = VT;
        v.xl[0] = j;
        v.ov = VX;
        v.ol = j;
        v.omap = j;
        return v;
    };
    if (m == 1) {
        STri t;
        t.line = 0;
        t.dd.line = 0;
        t.dd.pl[0] = 0;
        t.dd.pl[1] = 1;
        step_tri({t});
        SolveBlock b;
        SDg c;
        c.ov = VT; c.ol = 0; c.x0v = VX; c.x0l = 0;
        add_dg(b, {c});
        STrs ts;
        ts.tri = 0; ts.line = 0; ts.xv = VT; ts.ov = VX;
        add_ts(b, {ts});
        emit(b);
        P.nblocks = next_blk;
        return P;
    }
    std::map<std::pair<int, int>, SCoup> C;
    for (int u = 0; u < m; ++u) {
        if (m == 2) {
            SCoup c;
            c.d.line = u;
            c.d.pl[0] = 0;
            c.d.pl[1] = 1;
            C[{u, 1 - u}] = c;
        } else {
            SCoup cn, cp;
            cn.d.line = u;
            cn.d.pl[0] = 1;
            cp.d.line = u;
            cp.d.pl[0] = 0;
            C[{u, (u + 1) % m}] = cn;
            C[{u, (u + m - 1) % m}] = cp;
        }
    }
    std::vector<int> ring(m);
    std::iota(ring.begin(), ring.end(), 0);
    std::vector<SolveBlock> back;
    SolveBlock fwd;
    bool first = true;
    while (ring.size() > 1) {
        const int r = (int)ring.size();
        std::vector<char> el(r, 0);
        if (r == 2) el[1] = 1;
        else
            for (int p = 1; p + (r % 2) < r; p += 2) el[p] = 1;
        std::vector<int> elim, surv, nring;
        for (int p = 0; p < r; ++p) (el[p] ? elim : surv).push_back(p);
        for (int p : surv) nring.push_back(ring[p]);
        const int r2 = (int)nring.size();
        auto at = [&](int p) { return ring[((p % r) + r) % r]; };
        auto isel = [&](int p) { return el[((p % r) + r) % r] != 0; };
        std::vector<int> elines;
        for (int p : elim) elines.push_back(ring[p]);
        std::map<std::pair<int, int>, SCoup> NC;  // couplings of the new ring
        SolveBlock bk;
        if (first) {
            std::vector<STri> inv;
            for (int j : elines) {
                STri t;
                t.line = j;
                inv.push_back(t);
            }
            const int tri0 = (int)P.tris.size();  // tris index of elines[0]
            step_tri(inv);
            std::vector<SForm> post;
            auto add_term = [&](SForm& f, int j, const SDiag& rs, const SDiag& cs) {
                f.jl[f.nt] = j;
                f.rs[f.nt] = rs;
                f.cs[f.nt] = cs;
                ++f.nt;
            };
            for (int p : surv) {
                const int s = ring[p], right = at(p + 1), left = at(p - 1);
                const bool re = isel(p + 1), le = isel(p - 1);
                SForm f;
                f.out = s;
                f.own = s;
                if (re) add_term(f, right, C[{s, right}].d, C[{right, s}].d);
                if (le && left != right) add_term(f, left, C[{s, left}].d, C[{left, s}].d);
                post.push_back(f);
                if (r2 < 2) continue;
                const int nn = re ? at(p + 2) : right, np = le ? at(p - 2) : left;
                auto side = [&](SForm& g, bool e, int nb, int target) {
                    if (e) add_term(g, nb, C[{s, nb}].d, C[{nb, target}].d);
                    else g.dd = C[{s, nb}].d;
                };
                if (r2 == 2) {  // one combined coupling to the other survivor
                    SForm g;
                    g.out = alloc();
                    side(g, re, right, nn);
                    side(g, le, left, np);
                    SCoup c;
                    c.dense = true;
                    c.blk = g.out;
                    NC[{s, nn}] = c;
                    post.push_back(g);
                } else {
                    SForm gn, gp;
                    gn.out = alloc();
                    side(gn, re, right, nn);
                    gp.out = alloc();
                    side(gp, le, left, np);
                    SCoup cn, cp;
                    cn.dense = cp.dense = true;
                    cn.blk = gn.out;
                    cp.blk = gp.out;
                    NC[{s, nn}] = cn;
                    NC[{s, np}] = cp;
                    post.push_back(gn);
                    post.push_back(gp);
                }
            }
            step_form(post);
            // solve: t_j = X_j x_j (line solves) ; x_s -= C_sj .* t_j
            std::vector<STrs> ts;
            for (size_t q = 0; q < elines.size(); ++q) {
                STrs v;
                v.tri = tri0 + (int)q;
                v.line = elines[q];
                v.xv = VX;
                v.ov = VT;
                ts.push_back(v);
            }
            add_ts(fwd, ts);
            std::vector<SDg> dg;
            for (int p : surv) {
                const int s = ring[p], right = at(p + 1), left = at(p - 1);
                SDg d;
                d.ov = VX; d.ol = s; d.x0v = VX; d.x0l = s;
                if (isel(p + 1)) {
                    d.xv[d.nt] = VT; d.xl[d.nt] = right; d.d[d.nt] = C[{s, right}].d; ++d.nt;
                }
                if (isel(p - 1) && left != right) {
                    d.xv[d.nt] = VT; d.xl[d.nt] = left; d.d[d.nt] = C[{s, left}].d; ++d.nt;
                }
                dg.push_back(d);
            }
            add_dg(fwd, dg);
            // backward: t_j = x_j - C_ja .* x_a - C_jb .* x_b ; x_j = X_j t_j
            std::vector<SDg> bd;
            std::vector<STrs> bt;
            for (size_t q = 0; q < elim.size(); ++q) {
                const int p = elim[q];
                const int j = ring[p], a = at(p - 1), b = at(p + 1);
                SDg d;
                d.ov = VT; d.ol = j; d.x0v = VX; d.x0l = j;
                d.xv[0] = VX; d.xl[0] = a; d.d[0] = C[{j, a}].d; d.nt = 1;
                if (b != a) { d.xv[1] = VX; d.xl[1] = b; d.d[1] = C[{j, b}].d; d.nt = 2; }
                bd.push_back(d);
                STrs v;
                v.tri = tri0 + (int)q;
                v.line = j;
                v.xv = VT;
                v.ov = VX;
                bt.push_back(v);
            }
            add_dg(bk, bd);
            add_ts(bk, bt);
        } else {
            step_gj(elines);
            // Z(u, j) = C(u, j) X_j for each eliminated j and neighbour u
            std::map<std::pair<int, int>, int> Z;
            std::vector<SGemm> gz;
            std::vector<int> consumed;
            for (int p : elim) {
                const int j = ring[p];
                int nbs[2] = {at(p - 1), at(p + 1)};
                for (int q = 0; q < (nbs[0] == nbs[1] ? 1 : 2); ++q) {
                    const int u = nbs[q];
                    SGemm g;
                    g.out = alloc();
                    g.acc = 0;
                    g.sub = 0;
                    g.nt = 1;
                    g.A[0] = C[{u, j}].blk;
                    g.B[0] = j;
                    g.bmap[0] = j;
                    Z[{u, j}] = g.out;
                    gz.push_back(g);
                    consumed.push_back(C[{u, j}].blk);
                }
            }
            step_gemm(gz);
            for (int b : consumed) freel.push_back(b);
            std::vector<SGemm> gu;
            auto add = [&](SGemm& g, int zu, int bblk, int j) {
                g.A[g.nt] = zu;
                g.B[g.nt] = bblk;
                g.bmap[g.nt] = j;
                ++g.nt;
            };
            for (int p : surv) {
                const int s = ring[p], right = at(p + 1), left = at(p - 1);
                const bool re = isel(p + 1), le = isel(p - 1);
                SGemm d;
                d.out = s;
                d.acc = 1;
                if (re) add(d, Z[{s, right}], C[{right, s}].blk, right);
                if (le && left != right) add(d, Z[{s, left}], C[{left, s}].blk, left);
                gu.push_back(d);
                if (r2 < 2) continue;
                const int nn = re ? at(p + 2) : right, np = le ? at(p - 2) : left;
                if (r2 == 2) {
                    SGemm g;
                    int direct = -1;
                    if (!re) direct = C[{s, right}].blk;
                    if (!le) direct = C[{s, left}].blk;
                    if (direct >= 0) { g.out = direct; g.acc = 1; }
                    else { g.out = alloc(); g.acc = 0; }
                    if (re) add(g, Z[{s, right}], C[{right, nn}].blk, right);
                    if (le) add(g, Z[{s, left}], C[{left, np}].blk, left);
                    SCoup c;
                    c.dense = true;
                    c.blk = g.out;
                    NC[{s, nn}] = c;
                    gu.push_back(g);
                } else {
                    if (re) {
                        SGemm g;
                        g.out = alloc();
                        g.acc = 0;
                        add(g, Z[{s, right}], C[{right, nn}].blk, right);
                        SCoup c;
                        c.dense = true;
                        c.blk = g.out;
                        NC[{s, nn}] = c;
                        gu.push_back(g);
                    } else {
                        NC[{s, nn}] = C[{s, right}];
                    }
                    if (le) {
                        SGemm g;
                        g.out = alloc();
                        g.acc = 0;
                        add(g, Z[{s, left}], C[{left, np}].blk, left);
                        SCoup c;
                        c.dense = true;
                        c.blk = g.out;
                        NC[{s, np}] = c;
                        gu.push_back(g);
                    } else {
                        NC[{s, np}] = C[{s, left}];
                    }
                }
            }
            step_gemm(gu);
            // solve: x_s -= Z(s, j) x_j[prow_j] over the eliminated neighbours
            std::vector<SMv> mv;
            for (int p : surv) {
                const int s = ring[p], right = at(p + 1), left = at(p - 1);
                SMv v;
                v.bv = VX; v.bl = s; v.ov = VX; v.ol = s;
                if (isel(p + 1)) {
                    v.M[v.nt] = Z[{s, right}]; v.imap[v.nt] = right;
                    v.xv[v.nt] = VX; v.xl[v.nt] = right; ++v.nt;
                }
                if (isel(p - 1) && left != right) {
                    v.M[v.nt] = Z[{s, left}]; v.imap[v.nt] = left;
                    v.xv[v.nt] = VX; v.xl[v.nt] = left; ++v.nt;
                }
                mv.push_back(v);
            }
            add_mv(fwd, mv);
            // backward: t_j = x_j - C(j, a) x_a - C(j, b) x_b ; x_j = X_j t_j
            std::vector<SMv> b1, b2;
            for (int p : elim) {
                const int j = ring[p], a = at(p - 1), b = at(p + 1);
                SMv v;
                v.bv = VX; v.bl = j; v.ov = VT; v.ol = j;
                v.M[0] = C[{j, a}].blk; v.xv[0] = VX; v.xl[0] = a; v.nt = 1;
                if (b != a) { v.M[1] = C[{j, b}].blk; v.xv[1] = VX; v.xl[1] = b; v.nt = 2; }
                b1.push_back(v);
                b2.push_back(apply_x(j));
            }
            add_mv(bk, b1);
            add_mv(bk, b2);
        }
        back.push_back(bk);
        C.swap(NC);
        ring.swap(nring);
        first = false;
    }
    const int last = ring[0];
    step_gj({last});
    SDg c;
    c.ov = VT; c.ol = last; c.x0v = VX; c.x0l = last;
    add_dg(fwd, {c});
    add_mv(fwd, {apply_x(last)});
    emit(fwd);
    for (int i = (int)back.size() - 1; i >= 0; --i) emit(back[i]);
    P.nblocks = next_blk;
    return P;
}

}  // namespace bcr
}  // namespace fdfd
