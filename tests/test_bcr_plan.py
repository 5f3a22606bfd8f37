"""CPU checks of the host/device math headers of the reduction.

bcr_plan.h: tests/bcr_plan_check.cpp executes the symbolic plan with dense complex<double>
blocks, following the semantics of the device job kinds, on random periodic
5-point operators for line lengths n in {1,2,3,5,8} and ring sizes m in
{1..13, 16, 17, 23, 32, 33}.  The reduction is a direct method, so the solve
of A x = b must reach a relative residual below 1e-12; a plan that drops a
coupling, a corner block or a sign fails by orders of magnitude."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bcr_plan_is_a_direct_solver(tmp_path):
    exe = tmp_path / "bcr_plan_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fdfd.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "bcr_plan_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    lines = out.stdout.strip().splitlines()
    assert len(lines) == 90
    assert out.returncode == 0, "\n".join(l for l in lines if "FAIL" in l)


def test_tridiagonal_qr_inverse(tmp_path):
    """tests/tri_qr_check.cpp runs the Givens QR inverse of tri_qr.h (the
    exact code of the first-level device kernels) on random complex periodic
    tridiagonal blocks, n = 1..12, 17, 33, 64, 100, 257, diagonal dominant and
    indefinite: T X = I to 1e-12 relative.  A wrong rotation, a dropped wrap
    corner or a misplaced entry of R fails by orders of magnitude."""
    exe = tmp_path / "tri_qr_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fdfd.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "tri_qr_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    lines = out.stdout.strip().splitlines()
    assert len(lines) == 34
    assert out.returncode == 0, "\n".join(l for l in lines if "FAIL" in l)


def test_small_dense_eigensolver(tmp_path):
    """tests/small_eig_check.cpp runs small_eig.h (the projected-matrix
    eigensolver of both Arnoldi drivers, host and device) on random complex
    matrices (eigenpair residual <= 1e-12) and on matrices of prescribed
    spectrum, including a repeated eigenvalue (every eigenvalue to 1e-11)."""
    exe = tmp_path / "small_eig_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fdfd.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "small_eig_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    lines = out.stdout.strip().splitlines()
    assert len(lines) == 30
    assert out.returncode == 0, "\n".join(l for l in lines if "FAIL" in l)
