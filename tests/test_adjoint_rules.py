"""The adjoint's reverse-mode rules (nativeengine_amd/csrc/ne_adjoint.hpp)
against central finite differences of the forward step math, on the host:
quaternion rotation, inertia products, orientation integration, effective
mass and one Gauss-Seidel visit in each of its regimes (frictionless,
sliding, sticking, separating, against a static body)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reverse_rules_match_finite_differences(tmp_path):
    exe = str(tmp_path / "adjoint_rules")
    src = os.path.join(ROOT, "tests", "cpp", "adjoint_rules.cpp")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "include"), src, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "0 failing" in r.stdout
