"""The device query functions on the CPU.

merle_amd/csrc/devquery.hpp is compiled by g++ for the host against a
one-lane stand-in for the HIP runtime (tests/cpu_device), with
AddressSanitizer, and random programs (AND/OR/NOT/ANDM/ORM over exact and
edge bins, 1..300,007 rows) run through it: every evaluation form the
kernels use -- early-out, eager accumulators, straight into registers, and
the conjunction form, two- and three-valued -- must agree word for word,
and the fused and scan results must equal the oracle.  This pins the
evaluation logic (including the closed forms of the conjunction form) and
its bounds (tail words, early ends) without a GPU.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpu_device")


def _asan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                       text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


def test_device_functions_on_cpu(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    asan = _asan()
    lib = str(tmp_path / "libdevhost.so")
    flags = ["-O1", "-g", "-std=c++17", "-fPIC", "-shared"]
    if asan:
        flags += ["-fsanitize=address", "-fno-omit-frame-pointer"]
    cmd = (["g++"] + flags + ["-I" + SRC, "-I" + os.path.join(ROOT, "include"),
                              "-o", lib, os.path.join(SRC, "devhost.cpp"),
                              os.path.join(ROOT, "merle_amd", "csrc", "desc.cpp"),
                              os.path.join(ROOT, "merle_amd", "csrc", "pool.cpp")])
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ)
    if asan:
        env["LD_PRELOAD"] = asan
        env["ASAN_OPTIONS"] = "detect_leaks=0"
    r = subprocess.run([sys.executable, os.path.join(SRC, "run_devhost.py"), lib, "24"],
                       capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-3000:])
    last = r.stdout.strip().splitlines()[-1].split()
    assert int(last[-3]) > 0 and int(last[-1]) > 0   # both conjunction forms ran
