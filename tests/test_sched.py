"""The serving scheduler on the CPU.

merle_amd/csrc/sched.hpp (admission by memory, slot assignment, completion,
cancel) is compiled by g++ with AddressSanitizer and UBSan against a fake
device whose reservations come from the library's own host pool
(tests/cpu_sched/sched_host.cpp).  Scripted cases pin each rule -- among
them that a request the pool refuses while nothing is running or admitted
is rejected rather than waited on forever -- and random workloads (arrivals,
completions in any order, cancels, a fragmenting pool, a binding budget,
memory held by the caller) check after every pass that memory stays within
the budget and matches the reservations, that no request runs twice, and
that queued work never sits with every slot idle.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_scheduler_on_cpu(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "sched_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-omit-frame-pointer", "-I" + os.path.join(HERE, "cpu_device"),
           "-I" + os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(HERE, "cpu_sched", "sched_host.cpp"),
           os.path.join(ROOT, "merle_amd", "csrc", "pool.cpp")]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, "12"], capture_output=True, text=True, env=env,
                       timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert r.stdout.strip().splitlines()[-1].startswith("OK")
