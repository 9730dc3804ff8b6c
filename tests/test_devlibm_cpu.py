"""The library's device math (devlibm.hpp) against the host glibc, on the CPU.

devlibm.hpp is __host__ __device__ code whose every floating-point operation is explicit
(fma() where the glibc algorithm fuses, -ffp-contract=off elsewhere), so the host build
computes what the device computes.  Here it is compared with the host glibc on every one
of the 2^32 float inputs; test_devlibm_gpu.py repeats a large sample on the device.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("libm") / "devlibm_host_check")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "c++", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "ww3-gpu_amd", "csrc"),
           os.path.join(ROOT, "tests", "libm", "devlibm_host_check.cpp"), "-o", out, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.mark.parametrize("fn", ["expf"])
def test_every_float_input_matches_glibc(checker, fn):
    threads = str(max(1, min(8, os.cpu_count() or 1)))
    r = subprocess.run([checker, fn, threads], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout
