"""The batched form's host logic (job table, per-stream state, fork/join
stream capture, the autotune sweep) on a CPU stand-in of the HIP runtime
under AddressSanitizer (tests/hostsim/): no out-of-range launch, no
capture-rule violation, no invalid host or device access, for every batch
configuration the GPU tests use and the autotuner's sweep."""
import os
import pathlib
import shutil
import subprocess
import sys

import pytest

HERE = pathlib.Path(__file__).resolve().parent


def test_batch_host_logic_on_hip_standin(tmp_path):
    gxx = shutil.which("g++")
    if not gxx or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
        pytest.skip("g++ or the HIP headers are missing")
    subprocess.run(["bash", str(HERE / "hostsim" / "build.sh"), str(tmp_path)], check=True)
    asan = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True, check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, str(HERE / "hostsim" / "drive_batch.py"), str(tmp_path / "libomv_hostsim.so")],
                       capture_output=True, text=True, env=env, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "STUB-ERROR" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert "STUB_ERRORS 0" in out, out[-4000:]
