"""Builds and runs the C++ host-code tests (tests/cpp/) on the CPU."""
import os
import subprocess

from conftest import ROOT


def test_host_plan_and_residue_code(tmp_path):
    csrc = os.path.join(ROOT, "mersenne2gpu_amd", "csrc")
    exe = str(tmp_path / "test_host_plan")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17",
                    "-I", csrc, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_host_plan.cpp"), os.path.join(csrc, "m2_host.cpp"),
                    "-o", exe], check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]


def test_transform_arithmetic(tmp_path):
    """Split-twiddle products, lazy butterflies and the centre squaring formula
    (tests/cpp/test_arith.cpp) against plain modular arithmetic."""
    csrc = os.path.join(ROOT, "mersenne2gpu_amd", "csrc")
    exe = str(tmp_path / "test_arith")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", csrc, os.path.join(ROOT, "tests", "cpp", "test_arith.cpp"),
                    "-o", exe], check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]


def test_gerbicz_residue_arithmetic(tmp_path):
    """Addition, negation and division by 4 mod 2^p - 1 used by the Gerbicz
    check (tests/cpp/test_modn.cpp), against 128-bit integers and identities."""
    csrc = os.path.join(ROOT, "mersenne2gpu_amd", "csrc")
    exe = str(tmp_path / "test_modn")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", csrc, os.path.join(ROOT, "tests", "cpp", "test_modn.cpp"),
                    "-o", exe], check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
