"""Discrete-adjoint numerics checked on the CPU.

tests/adjoint_core_check.hip compiles the reverse functions of
mesoflow_amd/csrc/mf_adjoint.h and the forward device functions they reverse
(face_flux<P,5,D>, cutcell_values, chemistry + update_primitive_vars + clip,
externalbc, WENO5) for the host, and compares every reverse function with
central finite differences of its forward function along random directions.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_adjoint_reverse_functions_match_finite_differences(tmp_path):
    exe = str(tmp_path / "adjoint_core_check")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17", "-ffp-contract=off",
           "-Wno-unused-value", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "mesoflow_amd", "csrc"),
           os.path.join(ROOT, "tests", "adjoint_core_check.hip"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout[-4000:]
    # every family of checks ran
    for name in ("weno_plus", "weno_minus", "face D=0", "face D=1", "face D=2", "IB first order", "no hyperbolic part",
                 "post state", "post log_A", "externalbc inflow in_vel", "externalbc outflow", "cut cell",
                 "cut cell gather form"):
        assert name in r.stdout, name
