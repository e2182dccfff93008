"""The C++ COSMA interface without a GPU: headers, exported symbols, Strategy
validation and the native layout of CosmaMatrix.

The layout expectations follow src/cosma/mapper.cpp:105-200 for one rank: the
leaves of the sequential steps that split a matrix, depth first, each block
column-major with ld = its rows; intervals split as [len*i/d, len*(i+1)/d)
(src/cosma/interval.cpp:84-96).  Two layouts are written out by hand from
that rule; every other one is checked for the properties it implies (exact
tiling, contiguous offsets, storage order = depth-first order of the steps).
"""
import os
import re
import subprocess

import pytest

import cosma_amd as ca
from cosma_amd import build as cbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = cbuild.TEST_DRIVER


def driver():
    if not os.path.exists(DRIVER):
        cbuild.build_test_driver()
    return DRIVER


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def header_functions(name, pattern):
    text = open(os.path.join(ROOT, "include", "cosma", name)).read()
    return sorted(set(re.findall(pattern, text)))


def test_scalapack_symbols_all_spellings():
    """pxgemm.h declares the four spellings of p{s,d,c,z}gemm (src/cosma/pxgemm.h:6-107);
    the interposing library exports every one of them."""
    names = header_functions("pxgemm.h", r"void (p[sdcz]gemm_?|P[SDCZ]GEMM_?)\(")
    assert len(names) == 16
    syms = exported(cbuild.PXGEMM_LIB)
    assert set(names) <= syms, set(names) - syms


def test_prefixed_and_c_api_symbols_exported():
    names = header_functions("prefixed_pxgemm.h", r"void (cosma_p[sdcz]gemm_?|COSMA_P[SDCZ]GEMM_?)\(")
    names += header_functions("cinterface.hpp", r"void ([sdcz]multiply_using_layout)\(")
    assert len(names) == 20
    syms = exported(ca.LIB_PATH)
    assert set(names) <= syms, set(names) - syms
    # C++ entry points (mangled): local_multiply<double>, multiply<double>, pxgemm<double>
    demangled = subprocess.run(["nm", "-DC", "--defined-only", ca.LIB_PATH], capture_output=True, text=True,
                               check=True).stdout
    for sig in ("cosma::local_multiply<double>(cosma::cosma_context<double>*",
                "cosma::local_multiply_cpu<double>(double*, double*, double*, int, int, int, double, double)",
                "cosma::multiply<double>(cosma::CosmaMatrix<double>&",
                "cosma::pxgemm<double>(char, char, int, int, int, double",
                "cosma::get_context_instance<double>()",
                "cosma::Strategy::Strategy("):
        assert sig in demangled, sig


def run_strategy_driver():
    r = subprocess.run([driver(), "strategy"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def test_strategy_validation():
    lines = run_strategy_driver()
    throws = [l for l in lines if l.startswith("ok throws")]
    assert not [l for l in lines if l.startswith("FAIL")], lines
    assert len(throws) == 6
    assert any("P = 2 ranks" in l for l in throws)
    assert any("parallel steps need more than one rank" in l for l in throws)
    assert any("requires the memory for at least 300 units, but only 299" in l for l in throws)
    assert "ok memory_used 300" in lines      # m*k + k*n + m*n for 10^3


def parse_layouts(lines):
    out = {}
    for l in lines:
        if not l.startswith("layout "):
            continue
        parts = l.split()
        steps, label, shape, size = parts[1], parts[2], parts[3], int(parts[5])
        blocks = []
        for b in parts[6:]:
            dims, off = b.split("@")
            r0, nr, c0, nc = (int(x) for x in dims.split(":"))
            blocks.append((r0, nr, c0, nc, int(off)))
        rows, cols = (int(x) for x in shape.split("x"))
        out[(steps, label)] = (rows, cols, size, blocks)
    return out


def test_layouts_written_out():
    L = parse_layouts(run_strategy_driver())
    # m, n, k = 7, 9, 11.  "sm2,sk2": A (7 x 11) is split in m (3 | 4) then k (5 | 6),
    # leaves depth first: (m0,k0), (m0,k1), (m1,k0), (m1,k1).
    assert L[("sm2,sk2", "A")] == (7, 11, 77, [(0, 3, 0, 5, 0), (0, 3, 5, 6, 15), (3, 4, 0, 5, 33), (3, 4, 5, 6, 53)])
    # B (11 x 9) is not split by m: only k (5 | 6)
    assert L[("sm2,sk2", "B")] == (11, 9, 99, [(0, 5, 0, 9, 0), (5, 6, 0, 9, 45)])
    # "sn3,sm2": C split in n (3 | 3 | 3) first, then m
    assert L[("sn3,sm2", "C")][3] == [(0, 3, 0, 3, 0), (3, 4, 0, 3, 9), (0, 3, 3, 3, 21), (3, 4, 3, 3, 30),
                                     (0, 3, 6, 3, 42), (3, 4, 6, 3, 51)]
    # no steps: one column-major block
    for label, (r, c) in {"A": (7, 11), "B": (11, 9), "C": (7, 9)}.items():
        assert L[("-", label)] == (r, c, r * c, [(0, r, 0, c, 0)])


def test_layouts_tile_exactly():
    for (steps, label), (rows, cols, size, blocks) in parse_layouts(run_strategy_driver()).items():
        assert size == rows * cols
        cover = [[0] * cols for _ in range(rows)]
        off = 0
        for (r0, nr, c0, nc, o) in blocks:
            assert o == off, (steps, label)
            off += nr * nc
            for i in range(r0, r0 + nr):
                for j in range(c0, c0 + nc):
                    cover[i][j] += 1
        assert all(v == 1 for row in cover for v in row), (steps, label)


# ----------------------------------------------------------- environment knobs
KNOBS = {
    "COSMA_GPU_STREAMS": "5", "COSMA_GPU_MAX_TILE_M": "1000", "COSMA_GPU_MAX_TILE_N": "2000",
    "COSMA_GPU_MAX_TILE_K": "3000", "COSMA_GPU_MEMORY_PINNING": "off", "COSMA_CPU_MAX_MEMORY": "1",
    "COSMA_MEMORY_POOL_AMORTIZATION": "1.5", "COSMA_CPU_MEMORY_ALIGNMENT": "256", "COSMA_DIM_THRESHOLD": "77",
    "COSMA_ADAPT_STRATEGY": "OFF", "COSMA_OVERLAP_COMM_AND_COMP": "on", "COSMA_GPU_UNIFIED_MEMORY": "ON",
    "COSMA_MIN_LOCAL_DIMENSION": "33",
}


def run_driver(args, env_extra=None, drop_cosma=True, check=True):
    env = {k: v for k, v in os.environ.items() if not (drop_cosma and k.startswith("COSMA_"))}
    env.update(env_extra or {})
    r = subprocess.run([driver(), *args], capture_output=True, text=True, timeout=120, env=env)
    if check:
        assert r.returncode == 0, (r.stdout, r.stderr)
    return r


def env_values(env_extra=None):
    out = run_driver(["env"], env_extra).stdout
    return {l.split()[0]: l.split()[1:] for l in out.splitlines()}


def test_knob_defaults():
    """Defaults of src/cosma/environment_variables.hpp:10-97."""
    v = env_values()
    assert v["streams"] == ["2"] and v["tiles"] == ["5000", "5000", "5000"]
    assert v["cpu_max_memory"] == [str(2 ** 63 - 1)]
    assert v["pinning"] == ["1", "1"]               # pinned and the pool is page-locked
    assert v["amortization"] == ["1.2"] and v["alignment"] == ["0"] and v["dim_threshold"] == ["0"]
    assert v["adapt"] == ["1"] and v["overlap"] == ["0"] and v["unified"] == ["0"]
    assert v["min_local_dimension"] == ["200"]


def test_knobs_set():
    """Every knob read with the reference's parsing: numbers with atoi/atof,
    COSMA_CPU_MAX_MEMORY in MB converted to elements, booleans true only for
    "ON" in any case (src/cosma/environment_variables.cpp:9-21)."""
    v = env_values(KNOBS)
    assert v["streams"] == ["5"] and v["tiles"] == ["1000", "2000", "3000"]
    assert v["cpu_max_memory"] == [str(1024 * 1024 // 8)]
    assert v["pinning"] == ["0", "0"]               # "off" is false: pageable pool
    assert v["amortization"] == ["1.5"] and v["alignment"] == ["256"] and v["dim_threshold"] == ["77"]
    assert v["adapt"] == ["0"] and v["overlap"] == ["1"] and v["unified"] == ["1"]
    assert v["min_local_dimension"] == ["33"]
    # only "ON" (any case) is true
    assert env_values({"COSMA_GPU_MEMORY_PINNING": "1"})["pinning"] == ["0", "0"]
    assert env_values({"COSMA_ADAPT_STRATEGY": "yes"})["adapt"] == ["0"]
    assert env_values({"COSMA_OVERLAP_COMM_AND_COMP": "oN"})["overlap"] == ["1"]


def test_cpu_max_memory_limits_pdgemm():
    """pdgemm needs m*k + k*n + m*n elements on one rank; with
    COSMA_CPU_MAX_MEMORY (MB) below that it throws the reference's message
    (src/cosma/strategy.cpp:373-381) before touching the device."""
    r = run_driver(["pdgemm", "N", "N", "300", "300", "300", "1", "0", "1", "1", "1", "1", "1", "1", "/dev/null"],
                   {"COSMA_CPU_MAX_MEMORY": "1"}, check=False)
    assert r.returncode == 1
    assert "requires the memory for at least 270000 units, but only 131072 units are allowed" in r.stderr


def pool_lines(env_extra):
    out = run_driver(["pool"], env_extra).stdout.splitlines()
    assert not [l for l in out if l.startswith("FAIL")], out
    return {l.split()[0]: l.split()[1:] for l in out if not l.startswith("ok ")}


def test_memory_pool_amortization_and_release_order():
    """reserve() takes the buffers times COSMA_MEMORY_POOL_AMORTIZATION
    (src/cosma/memory_pool.cpp:116-150): 1504 elements x 1.2 = 1805; buffers
    are a stack, one released below the top goes when the top does; growth
    keeps the contents; CosmaMatrix objects take their storage from the
    context's pool and give it back."""
    v = pool_lines({"COSMA_GPU_MEMORY_PINNING": "OFF"})
    assert v["reserved"][:1] == ["1805"] and v["ids"] == ["0", "1001", "1501", "size", "1504"]
    assert v["matrices"] == ["1200", "800", "600", "pool", "2600"]
    assert v["pool"][-1] == "0"        # pageable storage with pinning off


def test_memory_pool_alignment():
    """COSMA_CPU_MEMORY_ALIGNMENT = 256 bytes: every buffer is padded to a
    multiple of 32 doubles, and the amortised reserve too."""
    v = pool_lines({"COSMA_GPU_MEMORY_PINNING": "OFF", "COSMA_CPU_MEMORY_ALIGNMENT": "256",
                    "COSMA_MEMORY_POOL_AMORTIZATION": "1.5"})
    assert v["reserved"] == ["2368", "padded1001", "1024", "sizes", "1024", "512", "32"]
    assert v["ids"] == ["0", "1024", "1536", "size", "1568"]
    assert v["matrices"] == ["1200", "800", "600", "pool", str(1216 + 800 + 608)]
