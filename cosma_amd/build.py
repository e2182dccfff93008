"""Builds the cosma_amd libraries.

  libcosma_amd.so         the gfx950 HIP kernels and the C ABI (csrc/*.hip,
                          compiled by hipcc) plus the C++ COSMA interface
                          (csrc/front/*.cpp except interpose.cpp, host code
                          compiled by g++);
  libcosma_amd_pxgemm.so  ScaLAPACK's p?gemm symbols (csrc/front/interpose.cpp),
                          linked against libcosma_amd.so.

Object files go to cosma_amd/_build/.  Run directly (`python -m cosma_amd.build`)
or through __graft_entry__.build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
FRONT = os.path.join(CSRC, "front")
LIB = os.path.join(HERE, "libcosma_amd.so")
PXGEMM_LIB = os.path.join(HERE, "libcosma_amd_pxgemm.so")
INTERPOSE = os.path.join(FRONT, "interpose.cpp")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"
# Packed-FP32 instructions (v_pk_*_f32) are disabled for all device code: some
# of their forms have produced wrong results on this hardware.
COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


HOST = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", FRONT]
# The C++ interface takes MPI_Comm: built against <mpi.h> from
# COSMA_AMD_MPI_INCLUDE when set (e.g. Open MPI, whose handles are pointers),
# else against the MPICH ABI's integer handles.
if os.environ.get("COSMA_AMD_MPI_INCLUDE"):
    HOST += ["-I", os.environ["COSMA_AMD_MPI_INCLUDE"]]


HOSTSRC = os.path.join(CSRC, "host")


def _host_sources():
    """Plain C++ sources of the library (compiled by g++): the COSMA
    interface (csrc/front) and the host-side helpers (csrc/host)."""
    out = [os.path.join(FRONT, n) for n in sorted(os.listdir(FRONT))
           if n.endswith(".cpp") and os.path.join(FRONT, n) != INTERPOSE]
    out += [os.path.join(HOSTSRC, n) for n in sorted(os.listdir(HOSTSRC)) if n.endswith(".cpp")]
    return out


def _sources():
    out = [os.path.join(CSRC, n) for n in sorted(os.listdir(CSRC)) if n.endswith(".hip")]
    return out + _host_sources()


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    else:
        cmd = [CXX, *HOST, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-ldl", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    tmp = PXGEMM_LIB + ".tmp"
    cmd = [CXX, *HOST, "-shared", "-o", tmp, INTERPOSE, "-L", HERE, "-l:libcosma_amd.so",
           "-Wl,-rpath,$ORIGIN", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PXGEMM_LIB)
    if verbose:
        print(f"built {LIB} and {PXGEMM_LIB}")
    return LIB


MINIAPP_SRC = os.path.join(HERE, "miniapp", "cosma_miniapp.cpp")
MINIAPP = os.path.join(HERE, "bin", "cosma_miniapp")


def build_miniapp(verbose=False):
    """Links cosma_amd/bin/cosma_miniapp (the cosma::multiply driver) against libcosma_amd.so."""
    os.makedirs(os.path.dirname(MINIAPP), exist_ok=True)
    tmp = MINIAPP + ".tmp"
    cmd = [CXX, *HOST, "-o", tmp, MINIAPP_SRC, "-L", HERE, "-l:libcosma_amd.so", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, MINIAPP)
    if verbose:
        print(f"built {MINIAPP}")
    return MINIAPP


TEST_DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "front_test.cpp")
TEST_DRIVER = os.path.join(ROOT, "tests", "cpp", "front_test")


def build_test_driver(verbose=False):
    """Links tests/cpp/front_test (the C++ interface's test driver) against libcosma_amd.so."""
    tmp = TEST_DRIVER + ".tmp"
    cmd = [CXX, *HOST, "-o", tmp, TEST_DRIVER_SRC, "-L", HERE, "-l:libcosma_amd.so",
           "-Wl,-rpath,$ORIGIN/../../cosma_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, TEST_DRIVER)
    if verbose:
        print(f"built {TEST_DRIVER}")
    return TEST_DRIVER


CPP_TESTS = os.path.join(ROOT, "tests", "cpp")
SANITIZERS = {"asan": "address,undefined", "tsan": "thread"}


def sanitized_exe(name, kind):
    return os.path.join(CPP_TESTS, f"{name}_{kind}")


def build_sanitized(verbose=False):
    """Host-only sanitizer builds (g++, no device code) of the C++ interface,
    the host-side helpers and the CPU oracle, driven by
    tests/cpp/front_test.cpp and tests/cpp/sanitize_test.cpp:
      front_test_asan      the COSMA interface (front/*.cpp, host/*.cpp)
                           under AddressSanitizer + UBSan, linked against
                           libcosma_amd.so for the C ABI it calls;
      sanitize_test_asan   the host packing (host/*.cpp) and the oracle
                           (oracle/oracle.c) under AddressSanitizer + UBSan;
      sanitize_test_tsan   the threaded host packing under ThreadSanitizer.
    The reference's own CI runs its tests in an AddressSanitizer image
    (docker/asan/deploy.Dockerfile); GPU code is never built with sanitizers."""
    out = {}
    common = ["-g", "-O1", "-fno-omit-frame-pointer", "-fno-sanitize-recover=all"]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", FRONT, "-I", CSRC]
    for kind in ("asan", "tsan"):
        san = ["-fsanitize=" + SANITIZERS[kind]]
        exe = sanitized_exe("sanitize_test", kind)
        cmd = [CXX, "-std=c++17", *common, *san, *inc, "-o", exe + ".tmp", os.path.join(CPP_TESTS, "sanitize_test.cpp"),
               *[os.path.join(HOSTSRC, n) for n in sorted(os.listdir(HOSTSRC)) if n.endswith(".cpp")],
               "-x", "c", os.path.join(ROOT, "oracle", "oracle.c"), "-x", "none", "-lpthread", "-lm"]
        _run(cmd)
        os.replace(exe + ".tmp", exe)
        out["sanitize_test_" + kind] = exe
    exe = sanitized_exe("front_test", "asan")
    cmd = [CXX, "-std=c++17", *common, "-fsanitize=" + SANITIZERS["asan"], *inc, "-o", exe + ".tmp",
           TEST_DRIVER_SRC, *_host_sources(), "-L", HERE, "-l:libcosma_amd.so", "-ldl", "-lpthread",
           "-Wl,-rpath,$ORIGIN/../../cosma_amd"]
    _run(cmd)
    os.replace(exe + ".tmp", exe)
    out["front_test_asan"] = exe
    if verbose:
        print(f"built the sanitizer drivers in {CPP_TESTS}")
    return out
STUB_LIB = os.path.join(CPP_TESTS, "libstub_scalapack.so")
# interposition test executables: name -> libraries in link (= lookup) order
INTERPOSE_LINKS = {
    "interpose_first": ["libcosma_amd_pxgemm.so", "libstub_scalapack.so"],
    "cosma_first": ["libcosma_amd.so", "libcosma_amd_pxgemm.so", "libstub_scalapack.so"],
    "stub_only": ["libstub_scalapack.so"],
    "no_stub": ["libcosma_amd_pxgemm.so"],
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_interpose_tests(verbose=False):
    """Builds the stand-in ScaLAPACK (tests/cpp/libstub_scalapack.so) and the
    interposition test driver linked in each order of INTERPOSE_LINKS."""
    _run([CXX, *HOST, "-shared", "-o", STUB_LIB + ".tmp", os.path.join(CPP_TESTS, "stub_scalapack.cpp")])
    os.replace(STUB_LIB + ".tmp", STUB_LIB)
    src = os.path.join(CPP_TESTS, "interpose_test.cpp")
    out = {}
    for name, libs in INTERPOSE_LINKS.items():
        exe = os.path.join(CPP_TESTS, "interpose_" + name)
        cmd = [CXX, *HOST, "-o", exe + ".tmp", src, "-L", HERE, "-L", CPP_TESTS, "-Wl,--no-as-needed"]
        cmd += ["-l:" + lib for lib in libs]
        cmd += ["-Wl,--as-needed", "-ldl", "-Wl,-rpath,$ORIGIN/../../cosma_amd:$ORIGIN"]
        _run(cmd)
        os.replace(exe + ".tmp", exe)
        out[name] = exe
    if verbose:
        print(f"built the interposition tests in {CPP_TESTS}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_miniapp(verbose=True)
    build_test_driver(verbose=True)
    build_interpose_tests(verbose=True)
    sys.exit(0)
