"""ScaLAPACK interposition on the CPU: libcosma_amd_pxgemm.so serves the p?gemm
symbols and passes what it does not run on to the next definition in the
process (src/cosma/pxgemm.cpp:40-71, interpose.h:33-51): pdgemm problems with
min(m, n, k) < COSMA_DIM_THRESHOLD and the s/c/z types.

A stand-in ScaLAPACK (tests/cpp/stub_scalapack.cpp) marks the calls it
receives.  The driver (tests/cpp/interpose_test.cpp) is linked in three
lookup orders and run once with the interposer preloaded; each run calls one
symbol on a 2 x 2 x 2 problem.
"""
import os
import subprocess

import pytest

from cosma_amd import build as cbuild

SPELLINGS = ["pdgemm_", "pdgemm", "PDGEMM_"]


@pytest.fixture(scope="module")
def exes():
    need = [os.path.join(cbuild.CPP_TESTS, "interpose_" + n) for n in cbuild.INTERPOSE_LINKS]
    if not all(os.path.exists(p) for p in need) or not os.path.exists(cbuild.STUB_LIB):
        return cbuild.build_interpose_tests()
    return {n: os.path.join(cbuild.CPP_TESTS, "interpose_" + n) for n in cbuild.INTERPOSE_LINKS}


def run(exe, call, threshold=None, preload=False):
    env = dict(os.environ)
    env.pop("COSMA_DIM_THRESHOLD", None)
    if threshold is not None:
        env["COSMA_DIM_THRESHOLD"] = str(threshold)
    if preload:
        # keep whatever the environment already preloads, add the interposer after it
        prev = env.get("LD_PRELOAD", "")
        env["LD_PRELOAD"] = (prev + ":" if prev else "") + cbuild.PXGEMM_LIB
    r = subprocess.run([exe, call], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (exe, call, r.stdout, r.stderr)
    return r.stdout.strip()


def cosma_took_it(out):
    """The interposer ran the call itself: on a machine without a GPU the
    device handle cannot be created; with one, C = A * I = A, so c[0] = 1."""
    return out.startswith("throws cosma_context: handle_create") or out == "returned 1"


@pytest.mark.parametrize("order", ["interpose_first", "cosma_first"])
@pytest.mark.parametrize("call", SPELLINGS)
def test_small_pdgemm_goes_to_next_definition(exes, order, call):
    """Below COSMA_DIM_THRESHOLD the next pdgemm takes the call, whether the
    interposer or libcosma_amd.so comes first in the lookup order (the lookup
    must start after the interposer, not after libcosma_amd.so)."""
    assert run(exes[order], call, threshold=100) == "stub 1 -7777"


@pytest.mark.parametrize("order", ["interpose_first", "cosma_first"])
def test_pdgemm_at_threshold_stays_with_cosma(exes, order):
    # min(m, n, k) = 2 is not below a threshold of 2, nor of the default 0
    assert cosma_took_it(run(exes[order], "pdgemm_", threshold=2))
    assert cosma_took_it(run(exes[order], "pdgemm_"))


@pytest.mark.parametrize("call", ["psgemm_", "pzgemm_"])
@pytest.mark.parametrize("order", ["interpose_first", "cosma_first"])
def test_other_types_go_to_next_definition(exes, order, call):
    assert run(exes[order], call) == "stub 1 -7777"


def test_preloaded_interposer_hands_off(exes):
    assert run(exes["stub_only"], "pdgemm_", threshold=100, preload=True) == "stub 1 -7777"
    assert run(exes["stub_only"], "psgemm_", preload=True) == "stub 1 -7777"
    assert cosma_took_it(run(exes["stub_only"], "pdgemm_", preload=True))


def test_without_next_definition(exes):
    """No ScaLAPACK after the interposer: s/c/z raise, small pdgemm runs here."""
    out = run(exes["no_stub"], "psgemm_")
    assert out.startswith("throws psgemm_: cosma_amd implements double precision only"), out
    out = run(exes["no_stub"], "pzgemm_")
    assert out.startswith("throws pzgemm_: cosma_amd implements double precision only"), out
    assert cosma_took_it(run(exes["no_stub"], "pdgemm_", threshold=100))
