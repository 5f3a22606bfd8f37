// Host runtime structures of libfdfd (C++; no types from here cross the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <string>
#include <vector>

#include "fdfd_common.h"

namespace fdfd {

using cd = std::complex<double>;

enum ProfId { P_TRAIL = 0, P_PANEL = 1, P_SCHUR = 2, P_SWEEP = 3, P_STENCIL = 4, P_ZBUILD = 5,
              P_COUNT = 6 };

struct Prof {
    bool enabled = false;
    struct Pending { hipEvent_t a, b; int id; double flops, bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double ms[P_COUNT] = {0};
    long long launches[P_COUNT] = {0};
    double flops[P_COUNT] = {0};
    double bytes[P_COUNT] = {0};
};

// Block-tridiagonal factors of one operator (pol, omega) on the handle's grid.
struct Factor {
    int id = 0, pol = 1, precision = 0;
    double omega = 0;
    int n = 0, m = 0, orient = 0;
    long si = 1, sj = 1;
    double scale = 1;       // A' = scale * A
    double2* coef = nullptr;  // 5 planes of N
    double2 *sxf = nullptr, *sxb = nullptr, *syf = nullptr, *syb = nullptr;
    void* Y = nullptr;        // m * n * n complex (double2 or float2)
    size_t elem = 16;
    int *rho = nullptr, *rhoinv = nullptr, *kap = nullptr, *kapinv = nullptr;
    double* pivstat = nullptr;  // 3 per line
    int k = 0;                  // middle line of the twisted factorization
    // device sweep tables: [transpose][phase] with phase 0 forward, 1 middle, 2 backward
    SweepJob* sweep_tab[2] = {nullptr, nullptr};
    hipGraphExec_t sweep_exec[2] = {nullptr, nullptr};
    int nfwd = 0, nbwd = 0;     // number of forward / backward steps
    std::vector<int> fwd_jobs, bwd_jobs;  // jobs per step
    std::vector<int> fwd_has_a, bwd_has_a;  // step has a chain-A job (listed first)
    double t_factor_ms = 0;
    double pivot_growth = 0;
    int singular_line = -1;
    int pivot_fallback = 0;   // 1 if the full-column pivot search was used
};

// captured launch sequence of one factorization pass for a shape
struct FactorGraph {
    int n = 0, m = 0, prec = 0, nb = 0, nfac = 1;
    bool full = false;
    hipGraphExec_t exec = nullptr;
    LineJob* lj_d = nullptr;
    SchurJob* sj_d = nullptr;
};

struct Handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;          // second elimination chain
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    GridDesc g{};
    bool has_grid = false, has_eps = false;
    double2* eps = nullptr;
    std::map<int, Factor*> factors;
    int next_id = 1;
    std::string err;
    int panel_width = 0;  // 0 = default per precision
    // factorization work (two chains)
    int work_n = 0, work_ch = 0;
    int *used = nullptr, *q = nullptr, *piv = nullptr, *win = nullptr, *verr = nullptr;
    void *Vbuf = nullptr, *LU = nullptr, *Zb = nullptr;
    bool full_pivot_search = false;
    bool use_graphs = true;            // FDFD_GRAPHS=0 disables graph replay
    std::vector<FactorGraph> graphs;
    // sweep / Krylov work
    long work_N = 0;
    double2 *r_in = nullptr, *g_mid = nullptr, *x_out = nullptr;  // sweep vectors
    double2 *V = nullptr, *Zk = nullptr;  // Krylov bases
    int kdim = 0;
    double2 *tmp = nullptr, *tmp2 = nullptr, *bvec = nullptr, *xvec = nullptr;
    double2* dpart = nullptr;   // multidot partials
    double* rpart = nullptr;    // norm partials
    double2* hdev = nullptr;    // small device vector of coefficients
    double2* tpart = nullptr;   // transposed-sweep partials
    int nblk = 0;
    Prof prof;
};

// error helper: record message, return code
int fail(Handle* h, int code, const std::string& msg);

// Copy ordered on the handle's stream, complete on return.  The handle's
// streams are non-blocking, so a plain hipMemcpy (null stream) would not wait
// for kernels still queued on them; every host<->device transfer of the
// library goes through here instead.
int stream_copy(Handle* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);

}  // namespace fdfd

struct fdfd_handle { fdfd::Handle impl; };
