"""Host code under sanitizers, on the CPU (no device code is ever built with
them): the COSMA interface (Strategy, CosmaMatrix layouts, the memory pool,
the environment knobs) and the host packing of the streaming GEMM under
AddressSanitizer + UndefinedBehaviorSanitizer, the threaded packing under
ThreadSanitizer, and the CPU oracle under AddressSanitizer + UBSan.  The
reference's CI does the same with its own tests in an AddressSanitizer image
(docker/asan/deploy.Dockerfile).  Any report makes the run exit non-zero
(-fno-sanitize-recover)."""
import os
import subprocess

import pytest

from cosma_amd import build as cbuild

MARKERS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "ERROR: LeakSanitizer")


@pytest.fixture(scope="module")
def exes():
    return cbuild.build_sanitized()


def run(exe, args, env_extra=None):
    env = {k: v for k, v in os.environ.items() if not k.startswith("COSMA_")}
    env.update(env_extra or {})
    # leak checking covers what the program allocates; the HIP runtime that
    # libcosma_amd.so loads keeps some allocations until exit by design
    env.setdefault("ASAN_OPTIONS", "detect_leaks=0:abort_on_error=0")
    env.setdefault("TSAN_OPTIONS", "halt_on_error=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, env=env)
    report = [m for m in MARKERS if m in r.stderr]
    assert r.returncode == 0 and not report, (exe, args, r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    assert "FAIL" not in r.stdout, r.stdout
    return r.stdout


def test_host_packing_asan(exes):
    out = run(exes["sanitize_test_asan"], ["host_copy"])
    assert out.count("ok host_copy_2d") == 7


def test_host_packing_threads_tsan(exes):
    out = run(exes["sanitize_test_tsan"], ["host_copy"])
    assert out.count("ok host_copy_2d") == 7


def test_oracle_asan(exes):
    out = run(exes["sanitize_test_asan"], ["oracle"])
    assert "ok block-cyclic gather on a 1 x 1 grid" in out


@pytest.mark.parametrize("args,env", [(["strategy"], {}), (["env"], {}),
                                      (["env"], {"COSMA_GPU_STREAMS": "4", "COSMA_CPU_MAX_MEMORY": "10"}),
                                      (["pool"], {"COSMA_GPU_MEMORY_PINNING": "OFF"}),
                                      (["pool"], {"COSMA_GPU_MEMORY_PINNING": "OFF",
                                                  "COSMA_CPU_MEMORY_ALIGNMENT": "256",
                                                  "COSMA_MEMORY_POOL_AMORTIZATION": "1.5"})])
def test_interface_asan(exes, args, env):
    run(exes["front_test_asan"], args, env)
