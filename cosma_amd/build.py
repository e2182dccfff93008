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


def _sources():
    out = [os.path.join(CSRC, n) for n in sorted(os.listdir(CSRC)) if n.endswith(".hip")]
    out += [os.path.join(FRONT, n) for n in sorted(os.listdir(FRONT))
            if n.endswith(".cpp") and os.path.join(FRONT, n) != INTERPOSE]
    return out


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
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-ldl"]
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
