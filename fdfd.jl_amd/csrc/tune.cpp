// Autotuner for the factorization's kernel variants.
//
// No counterpart in the reference (its solver choice is the FDFD_SOLVER
// switch of src/solver/solver.jl:10).  The variant knobs of a handle are:
//   gemm_b3     complex64 Schur GEMMs as fp16x3 split products (1) or FP32 MFMA (0)
//   trail_h3    complex64 Gauss-Jordan trailing updates as fp16x3 (1) or FP32 (0)
//   fuse_panels two Gauss-Jordan panels per trailing pass (1) or one (0)
//   gemm_waves  waves per workgroup of the FP32/FP64 GEMM (8 or 4)
//   gemm3m      three-multiplication complex product in the FP32/FP64 GEMM
// fdfd_autotune times a whole step (assemble + factor + refined solve of a
// point source to 1e-10) for every candidate on the handle's current grid
// and eps, keeps the fastest on the handle, and, when FDFD_TUNE_CACHE names a
// file, records it there.  Handles created later look that file up by
// (n, m, precision, pol) when they factor, unless switches were forced by
// the environment.  The file is plain text, one entry per line:
//   n m precision pol gemm_b3 trail_h3 fuse_panels gemm_waves gemm3m ms
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fdfd.h"
#include "runtime.h"

namespace fdfd {

namespace {

struct Knobs {
    int b3, trail, fuse, waves, g3m;
};

Knobs get_knobs(const Handle* h) {
    return Knobs{h->gemm_b3, h->trail_h3, h->fuse_panels, h->gemm_waves, h->gemm3m};
}
void set_knobs(Handle* h, const Knobs& k) {
    h->gemm_b3 = k.b3;
    h->trail_h3 = k.trail;
    h->fuse_panels = k.fuse;
    h->gemm_waves = k.waves;
    h->gemm3m = k.g3m;
}

struct Entry {
    int n, m, prec, pol;
    Knobs k;
    double ms;
};

std::vector<Entry> read_cache(const char* path) {
    std::vector<Entry> out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        Entry e{};
        if (is >> e.n >> e.m >> e.prec >> e.pol >> e.k.b3 >> e.k.trail >> e.k.fuse >> e.k.waves >>
            e.k.g3m >> e.ms)
            out.push_back(e);
    }
    return out;
}

int write_cache(const char* path, const std::vector<Entry>& es) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return -1;
    fprintf(f, "# libfdfd tuning cache: n m precision pol gemm_b3 trail_h3 fuse_panels gemm_waves "
               "gemm3m ms\n");
    for (const Entry& e : es)
        fprintf(f, "%d %d %d %d %d %d %d %d %d %.4f\n", e.n, e.m, e.prec, e.pol, e.k.b3, e.k.trail,
                e.k.fuse, e.k.waves, e.k.g3m, e.ms);
    fclose(f);
    return std::rename(tmp.c_str(), path);
}

std::vector<Knobs> candidates(int precision) {
    std::vector<Knobs> c;
    if (precision == FDFD_FP64) {
        for (int waves : {8, 4})
            for (int g3m : {1, 0})
                for (int fuse : {1, 0}) c.push_back(Knobs{1, 1, fuse, waves, g3m});
    } else {
        for (int b3 : {1, 0})
            for (int trail : {1, 0})
                for (int fuse : {1, 0}) {
                    if (b3) c.push_back(Knobs{b3, trail, fuse, 8, 1});
                    else
                        for (int waves : {8, 4}) c.push_back(Knobs{b3, trail, fuse, waves, 1});
                }
    }
    return c;
}

}  // namespace

// Apply a cached entry for (n, m, precision, pol) to the handle, once per
// factorization, unless the handle was tuned explicitly or its switches came
// from the environment.
void rt_tune_lookup(Handle* h, int n, int m, int precision, int pol) {
    if (h->tuned_explicitly || h->env_knobs) return;
    const char* path = std::getenv("FDFD_TUNE_CACHE");
    if (!path || !path[0]) return;
    for (const Entry& e : read_cache(path))
        if (e.n == n && e.m == m && e.prec == precision && e.pol == pol) {
            set_knobs(h, e.k);
            return;
        }
}

}  // namespace fdfd

using namespace fdfd;

extern "C" int fdfd_autotune(fdfd_handle* hh, int pol, double omega, int precision,
                             fdfd_stats* stats) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h) return FDFD_ERR_INVALID;
    if (!h->has_grid || !h->has_eps)
        return fail(h, FDFD_ERR_STATE, "grid and eps must be set before fdfd_autotune");
    if (pol != FDFD_TM && pol != FDFD_TE) return fail(h, FDFD_ERR_INVALID, "pol must be TM or TE");
    if (precision != FDFD_FP64 && precision != FDFD_FP32_IR)
        return fail(h, FDFD_ERR_INVALID, "precision must be FDFD_FP64 or FDFD_FP32_IR");
    if (!(omega > 0)) return fail(h, FDFD_ERR_INVALID, "omega must be > 0");
    (void)hipSetDevice(h->device);
    const GridDesc& g = h->g;
    const long N = (long)g.Nx * g.Ny;
    const int n = std::min(g.Nx, g.Ny), m = std::max(g.Nx, g.Ny);
    // the tuning source: a unit point source at the centre of the grid
    std::vector<fdfd_complex> src(N, fdfd_complex{0, 0});
    src[(long)(g.Ny / 2) * g.Nx + g.Nx / 2] = fdfd_complex{0, 1};
    double2* dsrc = nullptr;
    double2* dout = nullptr;
    if (hipMalloc(&dsrc, sizeof(double2) * N) != hipSuccess ||
        hipMalloc(&dout, sizeof(double2) * 3 * N) != hipSuccess) {
        if (dsrc) (void)hipFree(dsrc);
        return fail(h, FDFD_ERR_OOM, "autotune buffers");
    }
    int rc = stream_copy(h, dsrc, src.data(), sizeof(double2) * N, hipMemcpyHostToDevice);
    const Knobs start = get_knobs(h);
    Knobs best = start;
    double best_ms = 1e300;
    for (const Knobs& k : candidates(precision)) {
        if (rc) break;
        set_knobs(h, k);
        h->tuned_explicitly = true;  // no cache lookup while timing
        double t_best = 1e300;
        for (int rep = 0; rep < 2 && rc == FDFD_OK; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            int fid = 0;
            fdfd_stats st{};
            rc = fdfd_factor(hh, pol, omega, precision, &fid, &st);
            if (rc) break;
            rc = fdfd_solve_d(hh, fid, 1, dsrc, dout, 1e-10, &st);
            (void)fdfd_release_factor(hh, fid);
            if (rc) break;
            const double ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t0).count();
            t_best = std::min(t_best, ms);
        }
        if (rc == FDFD_ERR_NOCONV) {  // a variant that cannot reach the tolerance is skipped
            rc = FDFD_OK;
            continue;
        }
        if (t_best < best_ms) {
            best_ms = t_best;
            best = k;
        }
    }
    (void)hipFree(dsrc);
    (void)hipFree(dout);
    set_knobs(h, rc ? start : best);
    h->tuned_explicitly = rc == FDFD_OK;
    if (rc) return rc;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->t_factor_ms = best_ms;  // the winning step time (factor + refined solve)
        stats->singular_line = -1;
    }
    const char* path = std::getenv("FDFD_TUNE_CACHE");
    if (path && path[0]) {
        std::vector<Entry> es = read_cache(path);
        es.erase(std::remove_if(es.begin(), es.end(), [&](const Entry& e) {
                     return e.n == n && e.m == m && e.prec == precision && e.pol == pol;
                 }), es.end());
        es.push_back(Entry{n, m, precision, pol, best, best_ms});
        if (write_cache(path, es) != 0)
            return fail(h, FDFD_ERR_INVALID, std::string("cannot write the tuning cache ") + path);
    }
    return FDFD_OK;
}

extern "C" int fdfd_get_tuning(fdfd_handle* hh, int* knobs, int nknobs) {
    Handle* h = hh ? &hh->impl : nullptr;
    if (!h || !knobs || nknobs < 5) return FDFD_ERR_INVALID;
    const Knobs k = get_knobs(h);
    knobs[0] = k.b3;
    knobs[1] = k.trail;
    knobs[2] = k.fuse;
    knobs[3] = k.waves;
    knobs[4] = k.g3m;
    return FDFD_OK;
}
