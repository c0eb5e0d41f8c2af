"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer, on the CPU.

Four programs are built with g++/gcc -fsanitize=address,undefined
-fno-sanitize-recover=all and run, the last also with clang -fsanitize=thread;
each prints "ok" only if its own checks pass and the sanitizers report nothing
(a report aborts the program):
  * tests/host_cpp/test_host_logic.cpp: the server's arena allocator against a
    reference model (including exhaustion and coalescing), the admission
    decision, and the served operation sequence for taped and checkpointed
    requests;
  * tests/host_cpp/test_host_setup.cpp: descriptor validation with malformed
    descriptors, the C++ records / beam moments / beam profiles, layout
    packing;
  * tests/host_cpp/oracle_driver.c: every entry point of the C oracle;
  * tests/host_cpp/test_server_threads.cpp: the serving server's threads
    (worker, submitting, polling, cancelling and waiting clients, a statistics
    reader) with the HIP runtime and the kernel launchers replaced by
    tests/host_cpp/fake_hip.cpp, plus an injected launch failure and
    allocation failures at creation.
The GPU parts of the library cannot run under these sanitizers (GPU
AddressSanitizer is not available), so only host code is covered here.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pnepma_amd", "csrc")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
HIP_HOST = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                                reason="needs g++ and the ROCm headers")


def _run(cmd_build, exe):
    subprocess.check_call(cmd_build, cwd=ROOT)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    print(out.stdout[-2000:], out.stderr[-4000:])
    assert out.returncode == 0 and out.stdout.strip().endswith("ok")


def test_host_logic_sanitized(tmp_path):
    exe = str(tmp_path / "host_logic")
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "tests/host_cpp/test_host_logic.cpp", "-o", exe], exe)


def test_host_setup_sanitized(tmp_path):
    exe = str(tmp_path / "host_setup")
    srcs = [os.path.join(CSRC, f) for f in ("pnx_host.cpp", "pnx_setup.cpp", "pnx_coef.cpp")]
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-fconstexpr-ops-limit=1000000000", "-fconstexpr-loop-limit=10000000",
          "tests/host_cpp/test_host_setup.cpp", *srcs, "-o", exe], exe)


def test_oracle_sanitized(tmp_path):
    exe = str(tmp_path / "oracle_driver")
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("pnx_ref.c", "pnx_ref_material.c", "pnx_ref_records.c")]
    _run(["gcc", "-std=c99", *SAN, "tests/host_cpp/oracle_driver.c", *srcs, "-lm", "-o", exe], exe)


SERVER_SRCS = [os.path.join(CSRC, f) for f in ("pnx_server.cpp", "pnx_host.cpp", "pnx_setup.cpp", "pnx_coef.cpp")] + [
    "tests/host_cpp/fake_hip.cpp", "tests/host_cpp/test_server_threads.cpp"]
CLANG = "/opt/rocm/llvm/bin/clang++"


def test_server_threads_sanitized(tmp_path):
    exe = str(tmp_path / "server_asan")
    _run(["g++", "-std=c++17", *SAN, *HIP_HOST, "-fconstexpr-ops-limit=1000000000", "-fconstexpr-loop-limit=10000000",
          *SERVER_SRCS, "-lpthread", "-o", exe], exe)


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs the ROCm clang (its ThreadSanitizer runtime)")
def test_server_threads_thread_sanitizer(tmp_path):
    """clang's ThreadSanitizer (g++ 11's does not intercept the
    pthread_cond_clockwait that std::condition_variable::wait_for calls, and
    reports the waits as a double lock)."""
    exe = str(tmp_path / "server_tsan")
    subprocess.check_call([CLANG, "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-fno-omit-frame-pointer", *HIP_HOST,
                           "-fconstexpr-steps=100000000", *SERVER_SRCS, "-lpthread", "-o", exe], cwd=ROOT)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    for _ in range(3):  # the threads interleave differently on every run
        out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
        print(out.stdout[-2000:], out.stderr[-6000:])
        assert out.returncode == 0 and out.stdout.strip().endswith("ok")


def test_request_validation_refuses_malformed_descriptors():
    """pnx_request_device_bytes validates a request's descriptor on the host, as
    pnx_server_submit does before anything is staged."""
    import ctypes as C
    from pnepma_amd import _lib
    from pnepma_amd.solver import make_desc
    from pnepma_amd import workloads
    p, _, _, _ = workloads.raster2d(N=5, M=(16, 12), n_steps=20, n_beams=1)
    good = make_desc(p, 0)
    assert _lib.lib().pnx_request_device_bytes(C.byref(good)) > 0
    bad = []
    for field, value in (("N", 4), ("n_elements", 0), ("n_lines", 9), ("n_steps", 0), ("E_cutoff", 1e9)):
        d = make_desc(p, 0)
        setattr(d, field, value)
        bad.append(d)
    d = make_desc(p, 0)
    d.M[0] = 1  # a resolved axis after a singleton one
    bad.append(d)
    for d in bad:
        assert _lib.lib().pnx_request_device_bytes(C.byref(d)) < 0
        assert _lib.lib().pnx_last_error()
