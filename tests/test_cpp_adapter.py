"""The C++ drop-in (qes-public_amd/host/Solver_HIP.h) compiled against the
minimal re-declaration of the reference's Solver / WINDSGeneralData /
SolverFactory and driven through a QES-style time-step loop."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_solver_hip.cpp")
LIBDIR = os.path.join(ROOT, "qes-public_amd", "qesmc")
ORACLE_DIR = os.path.join(ROOT, "oracle", "build")


def build(out_dir):
    exe = os.path.join(out_dir, "test_solver_hip")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "qes-public_amd", "host"),
           "-I", os.path.join(ROOT, "qes-public_amd", "host", "qes_compat"),
           SRC, "-o", exe,
           "-L", LIBDIR, "-lqesmc", "-L", ORACLE_DIR, "-loracle_qesmc",
           f"-Wl,-rpath,{LIBDIR}", f"-Wl,-rpath,{ORACLE_DIR}"]
    subprocess.run(cmd, check=True)
    return exe


def test_adapter_compiles_and_links(qesmc, oracle, tmp_path):
    exe = build(str(tmp_path))
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_adapter_time_steps_match_oracle(qesmc, oracle, tmp_path):
    exe = build(str(tmp_path))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASS" in out.stdout
