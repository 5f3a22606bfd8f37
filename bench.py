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
KERNELS = ["trailing", "panel", "form", "solve", "stencil", "zbuild", "gemm"]


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


def cpu_baseline(n_target: int, sizes=(256, 512)):
    """SuperLU (scipy splu, COLAMD, single-threaded) on the same workload
    generator at the sample sizes; 1024^2 time extrapolated with the measured
    growth between the two samples."""
    import scipy.sparse.linalg as spla

    import fdfd_oracle as O

    times = []
    for n in sizes:
        d = O.random_pixel_device(n, n, 20, 1.55, 2)
        om = float(d.omega[0])
        t0 = time.perf_counter()
        A = O.assemble(d.grid, d.eps_r, om, O.TM)
        b = O.rhs(om, d.src)
        lu = spla.splu(A.tocsc(), permc_spec="COLAMD")
        x = lu.solve(b)
        t = time.perf_counter() - t0
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        times.append((n, t, res))
    (n1, t1, _), (n2, t2, r2) = times
    expo = math.log(t2 / t1) / math.log(n2 / n1)
    t_target = t2 * (n_target / n2) ** expo
    return {
        "value": 1.0 / t_target,
        "unit": f"TM solves/s at {n_target}^2",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle sparse LU (scipy SuperLU, COLAMD, 1 thread) of the same random-pixel "
                   f"TM workload at {n1}^2 ({t1:.2f} s) and {n2}^2 ({t2:.2f} s, residual "
                   f"{r2:.1e}); {n_target}^2 time {t_target:.1f} s extrapolated with the "
                   f"measured exponent {expo:.2f} in the side length"),
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
                    help="skip the mode-source and modulated-solve legs (headline only)")
    ap.add_argument("--precisions", default="mixed,fp64,mixed-fp32gemm",
                    help="mixed: complex64 factors with fp16x3 GEMM products (default mixed "
                         "path); mixed-fp32gemm: complex64 factors with FP32 MFMA products; "
                         "fp64: complex128 factors")
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
    from fdfd.core import Grid
    from fdfd.solve import PRECISIONS, Engine

    n = a.n
    d = O.random_pixel_device(n, n, 20, 1.55, 2)
    g = d.grid
    om = float(d.omega[0])
    N = n * n
    grid = Grid(0.02, g.Npml, (g.bounds[0][0], g.bounds[1][0]), (g.bounds[0][1], g.bounds[1][1]),
                g.L0)
    eng = Engine(0)
    eng.set_grid(grid)
    eng_x = None  # handle with FP32 GEMM products on the mixed path (read at handle creation)
    if "mixed-fp32gemm" in a.precisions.split(","):
        os.environ["FDFD_MIXED_GEMM"] = "fp32"
        eng_x = Engine(0)
        eng_x.set_grid(grid)
        del os.environ["FDFD_MIXED_GEMM"]
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
        e = eng_x if pname == "mixed-fp32gemm" else eng
        prec = PRECISIONS["mixed" if pname == "mixed-fp32gemm" else pname]
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
    # roofline of the dominant kernel: the tighter of its MFMA ceiling
    # (algorithmic flops / matrix peak) and its HBM ceiling (algorithmic
    # bytes / 8 TB/s) is its bound; frac = achieved / that peak
    if head == "mixed":
        if dom == "gemm":
            # each real product of the default mixed GEMM is formed from
            # three fp16 MFMA products, so its ceiling is the fp16 peak / 3
            mpeak, pnote = PEAK["fp16"] / 3, "fp16 dense MFMA peak / 3 (fp16x3 products)"
        else:
            mpeak, pnote = PEAK["fp32"], "FP32 dense MFMA peak"
    else:
        mpeak, pnote = PEAK["fp64"], "FP64 dense MFMA peak"
    per_launch_f = kd["flops"] / kd["launches"]
    per_launch_b = kd["bytes"] / kd["launches"]
    t_m = per_launch_f / mpeak
    t_b = per_launch_b / PEAK["hbm"]
    if per_launch_f > 0 and t_m >= t_b:
        achieved = per_launch_f / avg_s
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": mpeak / 1e12,
                "unit": "TFLOP/s", "frac": achieved / mpeak, "traffic": None, "kernel": dom,
                "peak_note": pnote}
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
        rt = "float" if head == "mixed" else "double"
        # kernel-name prefix of the dominant slot (the mixed path's GEMM is
        # the split-operand fp16x3 kernel)
        # (mixed lines with n % 4 == 0 use the fp16x3 row-strip trailing kernel)
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
    # whole-factorization MFMA figure: 8 n^3 m real flops per factorization
    roof["factor_tflops"] = 8.0 * n ** 3 * n / (R["factor_ms"] / 1e3) / 1e12

    cpu = None if a.no_cpu else cpu_baseline(n)
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
        "dtype": ("c64 factors (GEMM products fp16x3 on row-scaled split operands, FP32 accumulation) + c128 refinement"
                  if head == "mixed" else "c128"),
        "data": "synthetic (seeded random-pixel coupler, generated in-process)",
        "config": {
            "workload": (f"TM driven solve, {n}x{n} cells, dl=0.02 um, 20-cell SC-PML, 600x600 "
                         "design of 8x8 pixels (Bernoulli eps in {1,12.25}, seed 2), two guides, "
                         "25-cell line source, lambda=1.55 um"),
            "n": n, "tol": a.tol,
            "plain_fp64_solves_per_s": results.get("fp64", {}).get("solves_per_s"),
            "mixed_fp32_gemm_solves_per_s": results.get("mixed-fp32gemm", {}).get("solves_per_s"),
            "mixed_over_plain": (results["mixed"]["solves_per_s"] / results["fp64"]["solves_per_s"]
                                 if "mixed" in results and "fp64" in results else None),
            "factor_ms": {k: v["factor_ms"] for k, v in results.items()},
            "solve_ms": {k: v["solve_ms"] for k, v in results.items()},
            "gmres_its": {k: v["gmres_its"] for k, v in results.items()},
            "residual": {k: v["residual"] for k, v in results.items()},
            "pivot_fallback": {k: v["pivot_fallback"] for k, v in results.items()},
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if not a.no_legs:
        line["modes"] = mode_leg(eng, d, not a.no_cpu)
        line["modulated"] = modulated_leg(eng, not a.no_cpu)
    for p in (eps_d, src_d, out_d):
        hip.free(p)
    eng.close()
    if eng_x is not None:
        eng_x.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
