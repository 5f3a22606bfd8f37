"""Benchmark: TM solves/s at 1024^2 to 1e-10 relative residual, mixed
(complex64 block factors + FP64 FGMRES refinement) vs plain FP64 block factors.

One step = set eps (device-resident) -> assemble + factor (block cyclic
reduction of the line ring) on the GPU -> refined solve of one source ->
field recovery (Ez, Hx, Hy) into a device-resident output.  Inputs are in
HBM before the timed region starts.

Prints one JSON line (see the keys below).  The per-kernel roofline figures
come from HIP events recorded around every launch on the library's stream
during extra profiled steps (graph replay is off while profiling, so those
steps are not part of the timed region); the CPU baseline is the oracle's
sparse LU timed on bounded samples of the same workload generator.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fdfd.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

PEAK = {"hbm": 8.0e12, "fp64": 78.6e12, "fp32": 157.3e12, "fp16": 2.5e15}
KERNELS = ["trailing", "panel", "form", "solve", "stencil", "zbuild", "gemm", "assembly"]


class Hip:
    """Minimal HIP runtime access for device buffers (no torch needed)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.lib.hipMemcpyDtoD.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]

    def upload(self, a: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def alloc(self, nbytes: int) -> C.c_void_p:
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def download(self, p, a: np.ndarray):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0

    def free(self, p):
        self.lib.hipFree(p)

    def sync(self):
        self.lib.hipDeviceSynchronize()


def _cpu_count():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def cpu_baseline(n_target: int):
    """The reference's CPU path restated: assemble the same sparse A and b
    (driven.jl:21-36) and solve x = A \\ b with scipy's SuperLU (COLAMD, one
    thread; the reference calls UMFPACK through Julia's lu).  Timed on the
    host cores of this machine at config C1 (200 x 120) and at the metric's
    workload (n_target^2, the same random-pixel coupler as the GPU steps):
    assembly + factor + solve, one run each."""
    import scipy
    import scipy.sparse.linalg as spla

    import fdfd_oracle as O

    def timed(d):
        om = float(d.omega[0])
        t0 = time.perf_counter()
        A = O.assemble(d.grid, d.eps_r, om, O.TM)
        b = O.rhs(om, d.src)
        t1 = time.perf_counter()
        lu = spla.splu(A.tocsc(), permc_spec="COLAMD")
        t2 = time.perf_counter()
        x = lu.solve(b)
        t3 = time.perf_counter()
        res = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
        return {"assembly_s": t1 - t0, "factor_s": t2 - t1, "solve_s": t3 - t2,
                "total_s": t3 - t0, "residual": res}

    c1 = timed(O.config_c1())
    big = timed(O.random_pixel_device(n_target, n_target, 20, 1.55, 2))
    return {
        "value": 1.0 / big["total_s"],
        "unit": f"TM solves/s at {n_target}^2",
        "cores": 1,
        "kind": "port",
        "host_cpus_available": _cpu_count(),
        "library": f"scipy {scipy.__version__} SuperLU (splu, COLAMD), single-threaded",
        "sample": (f"one TM solve of the {n_target}^2 random-pixel coupler (assembly "
                   f"{big['assembly_s']:.1f} s + factor {big['factor_s']:.1f} s + solve "
                   f"{big['solve_s']:.2f} s, residual {big['residual']:.1e}) and one of config C1 "
                   f"200x120 ({c1['total_s'] * 1e3:.0f} ms)"),
        "c1": c1,
        "n_target": big,
    }


def mode_leg(eng, d, with_cpu: bool):
    """Mode-source eigenproblems (fdfd_eigenmode): the x-normal slice of 51
    cells through the input guide at the source column, at 64 wavelengths
    in [1.50, 1.60] um, in one batched call (one workgroup per slice).  The
    oracle's dense restatement of the same 64 problems is timed beside it."""
    from fdfd import modes
    eps = np.ascontiguousarray(d.eps_r[50, 512 - 25:512 + 26])
    lams = np.linspace(1.50, 1.60, 64)
    from fdfd.core import C0
    oms = [2 * math.pi * C0 / (l * 1e-6) for l in lams]
    slices = [eps] * len(oms)
    modes.eigenmode_batch(slices, 0.02, 1, oms, 3.0, 1, eng=eng)  # warm
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        beta, _, st = modes.eigenmode_batch(slices, 0.02, 1, oms, 3.0, 1, eng=eng)
    wall = (time.perf_counter() - t0) / reps
    out = {"workload": "64 TM slices of 51 cells (input guide, lambda 1.50-1.60 um), nev=1",
           "wall_ms_per_batch": wall * 1e3, "kernel_ms_per_batch": st["t_solve_ms"],
           "slices_per_s": len(oms) / wall, "op_applications": st["iterations"]}
    if with_cpu:
        import fdfd_oracle as O
        t0 = time.perf_counter()
        ob = [O.eigenmode_1d(eps, 0.02, om, O.TM, 3.0, 1)[0][0] for om in oms]
        tc = time.perf_counter() - t0
        out["cpu_oracle_ms_per_batch"] = tc * 1e3
        out["cpu_kind"] = "port (numpy dense eig of each slice, 1 process)"
        out["max_rel_diff_vs_oracle"] = float(np.max(np.abs(beta[:, 0] - np.array(ob)) /
                                                     np.abs(np.array(ob))))
    return out


def modulated_leg(eng, with_cpu: bool):
    """MF-FDFD solve (fdfd_modulated_solve) of the modulated waveguide the
    reference's tests use: 500 x 200 cells (dl = 0.01 um), one sideband each
    side, a TM mode source, eps 12.25 guide, deps = exp(i q x) on half the
    guide over 3 um.  GPU wall time per solve (factor + coupled FGMRES +
    fields) for both precisions; the oracle's direct sparse solve of the same
    coupled system is timed beside it."""
    from fdfd import core, modes
    from fdfd.modulated import ModulatedDevice, setup_deps, solve_modulated
    a, q, L = 0.2202, 2.9263, 5.0
    g = core.Grid(0.01, (15, 15), (0.0, L), (-1.0, 1.0))
    d = ModulatedDevice(g, 2 * math.pi * 1.939e14, 4.541e14, 1)
    core.setup_eps(d, lambda x, y: -a / 2 <= y <= a / 2, 12.25)
    setup_deps(d, lambda x, y: 1 <= x <= L - 1 and -a / 2 <= y <= 0,
               lambda x, y: np.exp(1j * q * x))
    modes.add_mode(d, modes.Mode(1, core.XHAT, 3.5, core.Point(0.2, 0.0), 4 * a))
    out = {"workload": "modulated waveguide 500x200, K=1 (3 sidebands), mode source, tol 1e-10"}
    res = {}
    for prec in ("mixed", "fp64"):
        solve_modulated(d, precision=prec, tol=1e-10, eng=eng)  # warm
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            f, st = solve_modulated(d, precision=prec, tol=1e-10, eng=eng, return_stats=True)
        wall = (time.perf_counter() - t0) / reps
        res[prec] = f[0]
        out[prec] = {"wall_ms": wall * 1e3, "factor_ms": st[0]["t_factor_ms"],
                     "solve_ms": st[0]["t_solve_ms"], "fgmres_its": st[0]["iterations"],
                     "residual": st[0]["residual"]}
    if with_cpu:
        import fdfd_oracle as O
        og = O.Grid.create(0.01, (15, 15), (0.0, L), (-1.0, 1.0))
        o = O.ModulatedDevice(og, float(d.omega[0]), d.Omega, 1)
        o.eps_r[:] = d.eps_r
        o.deps_r[:] = d.deps_r
        O.add_mode(o, O.Mode(1, O.XHAT, 3.5, (0.2, 0.0), 4 * a))
        t0 = time.perf_counter()
        fo = O.solve_modulated(o)[0]
        out["cpu_oracle_s"] = time.perf_counter() - t0
        out["cpu_kind"] = "port (scipy spsolve/SuperLU of the coupled system + refinement, 1 thread)"
        sc = max(np.linalg.norm(x.data) for x in fo)
        out["max_rel_diff_vs_oracle"] = {p: float(max(np.linalg.norm(a_.data - b_.data) / sc
                                                       for a_, b_ in zip(res[p], fo)))
                                         for p in res}
    return out


def _cgrid(g):
    from fdfd.core import Grid
    return Grid(g.dx, g.Npml, (g.bounds[0][0], g.bounds[1][0]), (g.bounds[0][1], g.bounds[1][1]),
                g.L0)


def te_leg(eng, hip, d, eps_d, src_d, out_d, steps: int):
    """TE driven solves of the metric's workload (1024^2 coupler, Hz line
    source), same step as the headline: device-resident eps/src -> assemble ->
    factor -> refined solve to 1e-10 -> fields."""
    from fdfd import _lib
    from fdfd.solve import PRECISIONS
    om = float(d.omega[0])
    lib, h = eng.lib, eng.h
    out = {"workload": "TE driven solve of the same 1024^2 coupler (Hz line source), tol 1e-10"}
    for pname in ("mixed", "fp64"):
        def step():
            st = _lib.fdfd_stats()
            fid = C.c_int()
            _lib.check(lib, h, lib.fdfd_set_eps_d(h, eps_d))
            _lib.check(lib, h, lib.fdfd_factor(h, 2, om, PRECISIONS[pname], C.byref(fid),
                                               C.byref(st)))
            _lib.check(lib, h, lib.fdfd_solve_d(h, fid.value, 1, src_d, out_d, 1e-10, C.byref(st)))
            _lib.check(lib, h, lib.fdfd_release_factor(h, fid.value))
            return st.as_dict()
        step()
        hip.sync()
        t0 = time.perf_counter()
        sts = [step() for _ in range(steps)]
        hip.sync()
        dt = (time.perf_counter() - t0) / steps
        out[pname] = {"solves_per_s": 1 / dt, "ms_per_step": dt * 1e3,
                      "gmres_its": int(np.mean([x["iterations"] for x in sts])),
                      "residual": max(x["residual"] for x in sts)}
    out["mixed_over_plain"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    return out


def c3_leg(eng):
    """Batched sweep: 3 grids (256^2, 384x512, 512^2) x 64 wavelengths x 16
    point sources = 192 factorizations and 3072 refined solves (mixed, to
    1e-10) in one fdfd_batch_run.  Wall time of the second of two identical
    calls (eps and sources uploaded inside it; results stay on the device)."""
    import fdfd_oracle as O
    jobs = O.config_c3()
    spec = [(_cgrid(dd.grid), 1, dd.eps_r, oms, srcs) for dd, oms, srcs in jobs]
    nsolve = sum(len(oms) * srcs.shape[2] for _, oms, srcs in jobs)
    # the first call also maps ~90 GB of fresh device memory (results and
    # factor slabs); it is reported as cold_wall_s, the second call is timed
    t0 = time.perf_counter()
    eng.batch_solve(spec, precision="mixed", tol=1e-10, fetch=[])
    cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, st = eng.batch_solve(spec, precision="mixed", tol=1e-10, fetch=[])
    wall = time.perf_counter() - t0
    return {"workload": "3 grids (256^2, 384x512, 512^2) x 64 wavelengths (1.50-1.60 um) x 16 "
                        "point sources, complex64 factors + FP64 FGMRES to 1e-10, one call",
            "solves": nsolve, "factorizations": 192, "wall_s": wall, "solves_per_s": nsolve / wall,
            "cold_wall_s": cold, "factor_ms": st["t_factor_ms"], "solve_ms": st["t_solve_ms"],
            "fgmres_its": st["iterations"], "worst_residual": st["residual"]}


def c4_leg(eng, reps: int = 3):
    """Adjoint gradients of J = |m^T Ez|^2 on the 1024 x 768 coupler with a
    256 x 256 continuous design: forward factor + refined solve + pullback
    (transposed FGMRES on the same factor + gradient kernel) per gradient."""
    import fdfd_oracle as O
    from fdfd import adjoint, core
    d, m, cells, _ = O.config_c4()
    dev = core.Device(_cgrid(d.grid), float(d.omega[0]))
    dev.eps_r[:] = d.eps_r
    dev.src[:] = d.src
    mf = m.reshape(d.grid.N, order="F")
    out = {"workload": "dJ/deps, 1024x768, 65,536 design cells, tol 1e-10 (host arrays in/out)"}
    grads = {}
    for pname in ("mixed", "fp64"):
        adjoint.objective_and_gradient(dev, mf, precision=pname, tol=1e-10, eng=eng)  # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            J, g, _ = adjoint.objective_and_gradient(dev, mf, precision=pname, tol=1e-10, eng=eng)
        dt = (time.perf_counter() - t0) / reps
        grads[pname] = g.ravel(order="F")[cells]
        out[pname] = {"gradients_per_s": 1 / dt, "ms_per_gradient": dt * 1e3}
    out["mixed_vs_plain_rel"] = float(np.linalg.norm(grads["mixed"] - grads["fp64"]) /
                                      np.linalg.norm(grads["fp64"]))
    return out


def c5_leg(eng):
    """8 eigenfrequencies nearest 2 pi c0/1.55 um of the 768^2 triangular rod
    lattice with a missing rod (shift-invert Arnoldi on the factors, tol
    1e-10), plain and mixed factors; omega and Q differences between them."""
    import fdfd_oracle as O
    d = O.config_c5()
    eng.set_grid(_cgrid(d.grid))
    eng.set_eps(d.eps_r)
    out = {"workload": "TM eigenfrequencies, 768^2 rod lattice, nev=8, ncv=20, tol 1e-10"}
    ws = {}
    for pname in ("mixed", "fp64"):
        # untimed first call: device buffers for the new grid size
        eng.eigenfrequency(1, float(d.omega[0]), 8, ncv=20, precision=pname, tol=1e-10,
                           want_fields=False)
        t0 = time.perf_counter()
        w, _, st = eng.eigenfrequency(1, float(d.omega[0]), 8, ncv=20, precision=pname, tol=1e-10,
                                      want_fields=False)
        dt = time.perf_counter() - t0
        ws[pname] = w
        out[pname] = {"seconds": dt, "factor_ms": st["t_factor_ms"], "fgmres_its": st["iterations"],
                      "restarts": st["restarts"]}
    wp = ws["fp64"]
    wm = np.array([ws["mixed"][int(np.argmin(np.abs(ws["mixed"] - x)))] for x in wp])
    qp, qm = wp.real / (2 * np.abs(wp.imag)), wm.real / (2 * np.abs(wm.imag))
    out["omega_rel_diff"] = float(np.max(np.abs(wm - wp) / np.abs(wp)))
    out["Q_rel_diff"] = float(np.max(np.abs(qm - qp) / qp))
    out["Q"] = [float(q) for q in qp]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--profile-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip every leg but the headline TM solves (profiling runs)")
    ap.add_argument("--precisions", default="mixed,fp64,mixed-fp32gemm,mixed-fp32",
                    help="mixed: complex64 factors with fp16x3 GEMM and trailing products (default "
                         "mixed path); mixed-fp32gemm: FP32 MFMA GEMMs, fp16x3 trailing updates; "
                         "mixed-fp32: FP32 MFMA GEMMs and trailing updates; fp64: complex128 "
                         "factors")
    ap.add_argument("--probe-copy", type=int, default=0, metavar="MIB",
                    help="only run 4 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if a.probe_copy:
        hip = Hip()
        nbytes = a.probe_copy << 20
        src, dst = hip.alloc(nbytes), hip.alloc(nbytes)
        for _ in range(4):
            assert hip.lib.hipMemcpyDtoD(dst, src, nbytes) == 0
        hip.sync()
        hip.free(src)
        hip.free(dst)
        print(json.dumps({"probe": "copy", "bytes_read_per_copy": nbytes,
                          "bytes_written_per_copy": nbytes, "copies": 4}))
        return

    import fdfd_oracle as O
    from fdfd import _lib
    from fdfd.solve import PRECISIONS, Engine

    n = a.n
    d = O.random_pixel_device(n, n, 20, 1.55, 2)
    om = float(d.omega[0])
    N = n * n
    grid = _cgrid(d.grid)
    eng = Engine(0)
    eng.set_grid(grid)
    # handles whose mixed path uses FP32 MFMA products (switches read at creation)
    variant_env = {"mixed-fp32gemm": {"FDFD_MIXED_GEMM": "fp32"},
                   "mixed-fp32": {"FDFD_MIXED_GEMM": "fp32", "FDFD_MIXED_TRAIL": "fp32"}}
    engines = {}
    for pname in a.precisions.split(","):
        if pname in variant_env:
            os.environ.update(variant_env[pname])
            engines[pname] = Engine(0)
            engines[pname].set_grid(grid)
            for k in variant_env[pname]:
                del os.environ[k]
        else:
            engines[pname] = eng
    hip = Hip()
    eps_d = hip.upload(np.ascontiguousarray(d.eps_r.ravel(order="F")))
    src_d = hip.upload(np.ascontiguousarray(d.src.ravel(order="F")))
    out_d = hip.alloc(16 * 3 * N)

    def step(prec: int, e):
        lib, h = e.lib, e.h
        st = _lib.fdfd_stats()
        fid = C.c_int()
        _lib.check(lib, h, lib.fdfd_set_eps_d(h, eps_d))
        _lib.check(lib, h, lib.fdfd_factor(h, 1, om, prec, C.byref(fid), C.byref(st)))
        t_f = st.t_factor_ms
        _lib.check(lib, h, lib.fdfd_solve_d(h, fid.value, 1, src_d, out_d, a.tol, C.byref(st)))
        _lib.check(lib, h, lib.fdfd_release_factor(h, fid.value))
        r = st.as_dict()
        r["t_factor_ms"] = t_f
        return r

    results = {}
    for pname in a.precisions.split(","):
        e = engines[pname]
        prec = PRECISIONS["mixed" if pname.startswith("mixed") else pname]
        for _ in range(a.warmup):
            step(prec, e)
        hip.sync()
        t0 = time.perf_counter()
        stats = [step(prec, e) for _ in range(a.steps)]
        hip.sync()
        dt = time.perf_counter() - t0
        worst = max(s["residual"] for s in stats)
        if not worst <= a.tol:
            raise SystemExit(f"{pname}: residual {worst:.2e} above {a.tol:.0e}: result invalid")
        # profiled steps: per-kernel device time from HIP events on the stream
        e.profile(True)
        for _ in range(a.profile_steps):
            step(prec, e)
        kern = {k: e.profile_read(i) for i, k in enumerate(KERNELS)}
        e.profile(False)
        results[pname] = {
            "solves_per_s": a.steps / dt, "ms_per_step": 1e3 * dt / a.steps,
            "factor_ms": float(np.mean([s["t_factor_ms"] for s in stats])),
            "solve_ms": float(np.mean([s["t_solve_ms"] for s in stats])),
            "gmres_its": int(np.mean([s["iterations"] for s in stats])),
            "residual": worst, "pivot_fallback": max(s["pivot_fallback"] for s in stats),
            "kernels": kern,
        }
        print(f"# {pname}: {results[pname]['solves_per_s']:.3f} solves/s, "
              f"{results[pname]['ms_per_step']:.1f} ms/step (factor "
              f"{results[pname]['factor_ms']:.1f} ms, solve {results[pname]['solve_ms']:.1f} ms,"
              f" {results[pname]['gmres_its']} GMRES its, residual {worst:.1e})",
              file=sys.stderr, flush=True)
        for k, v in kern.items():
            if v["launches"]:
                print(f"#   {k:9s} {v['ms']:9.2f} ms / {v['launches']:6d} launches "
                      f"= {1e3 * v['ms'] / v['launches']:8.2f} us avg", file=sys.stderr)

    head = "mixed" if "mixed" in results else next(iter(results))
    R = results[head]
    # dominant kernel of the headline path, with its roofline
    dom = max(R["kernels"], key=lambda k: R["kernels"][k]["ms"])
    kd = R["kernels"][dom]
    if not kd["launches"]:  # --profile-steps 0: no per-kernel events were recorded
        kd = {"ms": 0.0, "launches": 1, "flops": 0.0, "bytes": 0.0}
    avg_s = kd["ms"] / kd["launches"] / 1e3
    # roofline of the dominant kernel against the published peak of the MFMA
    # instructions it runs (fp16 for the split-operand GEMM, FP32 or FP64
    # otherwise) or of HBM; achieved = algorithmic flops (8 n^3 per complex
    # GEMM) or bytes per launch / average launch time
    if head == "mixed":
        ptype = "fp16" if (dom == "gemm" or dom == "trailing") else "fp32"
    elif head.startswith("mixed"):
        ptype = "fp32"
    else:
        ptype = "fp64"
    mpeak = PEAK[ptype]
    # executed MFMA work per algorithmic flop: fp16x3 forms each real product
    # from three fp16 products
    exec_factor = 3.0 if ptype == "fp16" else 1.0
    per_launch_f = kd["flops"] / kd["launches"]
    per_launch_b = kd["bytes"] / kd["launches"]
    t_m = per_launch_f / mpeak
    t_b = per_launch_b / PEAK["hbm"]
    if per_launch_f > 0 and t_m * exec_factor >= t_b:
        achieved = per_launch_f / avg_s
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": mpeak / 1e12,
                "unit": "TFLOP/s", "frac": achieved / mpeak, "traffic": None, "kernel": dom,
                "peak_note": f"published dense {ptype} MFMA peak of MI355X",
                "mfma_executed_tflops": exec_factor * achieved / 1e12,
                "mfma_executed_frac": exec_factor * achieved / mpeak,
                "executed_note": ("fp16x3 split products: 3 fp16 MFMA products per real "
                                  "product" if exec_factor > 1 else "executed = algorithmic")}
        if per_launch_b > 0:
            roof["hbm_achieved_GBps"] = per_launch_b / avg_s / 1e9
            roof["flop_per_byte"] = per_launch_f / per_launch_b
    elif per_launch_b > 0:
        achieved = per_launch_b / avg_s
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9,
                "unit": "GB/s", "frac": achieved / PEAK["hbm"], "traffic": None, "kernel": dom}
    else:
        roof = {"bound": "latency", "achieved": None, "peak": None, "unit": None, "frac": None,
                "traffic": None, "kernel": dom,
                "note": "dominant kernel has no algorithmic flop/byte count (pivot search)"}
    roof["launches"] = kd["launches"]
    # HBM traffic per launch of the dominant kernel from the committed PMC
    # summary (FETCH_SIZE / WRITE_SIZE runs scaled by a copy of known size)
    pmc = os.path.join(ROOT, "profiles", f"pmc_traffic_{n}_tm_{head}.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            P = json.load(fh)
        rt = "float" if head.startswith("mixed") else "double"
        trail = "k_trailing_h3" if head == "mixed" and n % 4 == 0 else "k_trailing<" + rt
        prefix = {"gemm": "k_gemm_b3" if head == "mixed" else "k_gemm<" + rt,
                  "trailing": trail, "solve": "k_mv<" + rt}.get(dom)
        for name, v in P["kernels"].items():
            if prefix and name.startswith(prefix):
                roof["traffic"] = v["read_bytes"] + v["write_bytes"]
                roof["traffic_unit"] = "bytes/launch"
                roof["traffic_source"] = os.path.relpath(pmc, ROOT)
                break
    roof["avg_launch_us"] = avg_s * 1e6
    # whole-factorization figure: F = 8 n^3 m real flops per factorization
    roof["factor_tflops"] = 8.0 * n ** 3 * n / (R["factor_ms"] / 1e3) / 1e12
    roof["factor_frac_fp32_peak"] = roof["factor_tflops"] * 1e12 / PEAK["fp32"]
    # the stencil / residual kernel (matrix-free operator apply, 64 B/cell)
    ks = R["kernels"]["stencil"]
    stencil = None
    if ks["launches"] and ks["bytes"] > 0:
        gbps = ks["bytes"] / (ks["ms"] / 1e3) / 1e9
        stencil = {"kernel": "k_apply (residual b - A'x / A'x)", "algorithmic_bytes_per_cell": 64,
                   "launches": ks["launches"], "avg_launch_us": 1e3 * ks["ms"] / ks["launches"],
                   "achieved_GBps": gbps, "frac_hbm": gbps * 1e9 / PEAK["hbm"]}

    legs = {}
    if not a.no_legs:
        print("# TE leg", file=sys.stderr, flush=True)
        legs["te"] = te_leg(eng, hip, d, eps_d, src_d, out_d, max(2, a.steps // 4))
    for p_ in (eps_d, src_d, out_d):
        hip.free(p_)
    if not a.no_legs:
        print("# C4 leg", file=sys.stderr, flush=True)
        legs["c4_adjoint"] = c4_leg(eng)
        print("# C5 leg", file=sys.stderr, flush=True)
        legs["c5_eigen"] = c5_leg(eng)
        print("# mode and modulated legs", file=sys.stderr, flush=True)
        eng.set_grid(grid)
        legs["modes"] = mode_leg(eng, d, not a.no_cpu)
        legs["modulated"] = modulated_leg(eng, not a.no_cpu)
        # last: the batch frees ~26 GB of results when it ends, and the
        # driver's clearing of freed memory slows the GPU work right after
        print("# C3 leg", file=sys.stderr, flush=True)
        legs["c3_batch"] = c3_leg(eng)
    cpu = None
    if not a.no_cpu:
        print("# CPU baseline (SuperLU at C1 and the metric size)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n)
    line = {
        "metric": "TM solves/s at 1024^2 to 1e-10 rel. residual, mixed vs plain FP64 GPU path",
        "value": R["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64 factors (GEMM and trailing products fp16x3 on row-scaled split operands, "
                  "FP32 accumulation) + c128 refinement" if head == "mixed" else "c128"),
        "data": "synthetic (seeded random-pixel coupler, generated in-process)",
        "config": {
            "workload": (f"TM driven solve, {n}x{n} cells, dl=0.02 um, 20-cell SC-PML, 600x600 "
                         "design of 8x8 pixels (Bernoulli eps in {1,12.25}, seed 2), two guides, "
                         "25-cell line source, lambda=1.55 um"),
            "n": n, "tol": a.tol,
            "plain_fp64_solves_per_s": results.get("fp64", {}).get("solves_per_s"),
            "mixed_fp32_gemm_solves_per_s": results.get("mixed-fp32gemm", {}).get("solves_per_s"),
            "mixed_fp32_solves_per_s": results.get("mixed-fp32", {}).get("solves_per_s"),
            "mixed_over_plain": (results["mixed"]["solves_per_s"] / results["fp64"]["solves_per_s"]
                                 if "mixed" in results and "fp64" in results else None),
            "factor_ms": {k: v["factor_ms"] for k, v in results.items()},
            "solve_ms": {k: v["solve_ms"] for k, v in results.items()},
            "gmres_its": {k: v["gmres_its"] for k, v in results.items()},
            "residual": {k: v["residual"] for k, v in results.items()},
            "pivot_fallback": {k: v["pivot_fallback"] for k, v in results.items()},
        },
        "roofline": roof,
        "stencil": stencil,
        "cpu_baseline": cpu,
    }
    line.update(legs)
    eng.close()
    for pname, e in engines.items():
        if e is not eng:
            e.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
