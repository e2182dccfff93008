"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/cosma_amd.h, and rejects bad arguments with a
status code and an error text (the C++ layer turns these into
std::runtime_error, like the reference's check_runtime_status,
src/cosma/gpu/utils.hpp:7-13)."""
import ctypes
import os
import re
import subprocess

import pytest

import cosma_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cosma_amd_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads_and_reports_version():
    assert os.path.exists(ca.LIB_PATH), "build first: python -m cosma_amd.build"
    assert ca.lib().cosma_amd_version().decode().startswith("cosma_amd")


def test_every_declared_entry_point_is_exported():
    decl = declared_functions()
    assert len(decl) >= 6
    syms = exported_symbols(ca.LIB_PATH)
    missing = [d for d in decl if d not in syms]
    assert not missing, f"declared but not exported: {missing}"
    # and the Python binding knows every one of them
    assert set(decl) <= set(ca.SIGNATURES), set(decl) - set(ca.SIGNATURES)


def test_null_handle_is_rejected_without_touching_the_gpu():
    L = ca.lib()
    rc = L.cosma_amd_dgemm(None, b"N", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, None)
    assert rc == ca.ERR_INVALID_ARG
    assert "null handle" in L.cosma_amd_last_error().decode()
    rc = L.cosma_amd_block_transform(None, b"T", 4, 4, 1.0, None, 4, 0.0, None, 4, None)
    assert rc == ca.ERR_INVALID_ARG


def test_bad_arguments_raise_with_text():
    L = ca.lib()
    with pytest.raises(ca.CosmaAmdError, match="negative extent"):
        ca.check(L.cosma_amd_fill_uniform(None, -1, 4, 4, 1, 0, None))
    with pytest.raises(ca.CosmaAmdError, match="ld="):
        ca.check(L.cosma_amd_fill_uniform(ctypes.c_void_p(16), 8, 4, 4, 1, 0, None))
    with pytest.raises(ca.CosmaAmdError, match="tile_config"):
        ca.set_dgemm_tuning(7, 0)
    # empty problems are valid no-ops
    assert L.cosma_amd_fill_uniform(None, 0, 4, 1, 1, 0, None) == ca.OK


def test_header_is_plain_c():
    # The boundary header must compile as C (no C++ or torch types).
    src = "#include \"cosma_amd.h\"\nint main(void){return 0;}\n"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-x", "c",
                        "-fsyntax-only", "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_python_constants_match_the_header():
    # The enum values the Python binding passes through the ABI are the ones
    # the header declares (compiled and printed by the C compiler).
    names = ["COSMA_AMD_OZAKI_MODULI", "COSMA_AMD_OZAKI_SLICES", "COSMA_AMD_OZAKI_MODULI_BF16",
             "COSMA_AMD_OZAKI_STRICT", "COSMA_AMD_OZAKI_SPLIT_MIN_DIM"]
    src = "#include <stdio.h>\n#include \"cosma_amd.h\"\nint main(void){" + "".join(
        f'printf("%d\\n", (int){n});' for n in names) + "return 0;}\n"
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "enum")
        r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-x", "c", "-", "-o", exe],
                           input=src, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        values = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert values == [ca.OZAKI_MODULI, ca.OZAKI_SLICES, ca.OZAKI_MODULI_BF16, ca.OZAKI_STRICT,
                      ca.OZAKI_SPLIT_MIN_DIM]


def test_info_struct_layout_matches_the_header():
    # size and field offsets of cosma_amd_ozaki_info as the C compiler lays it out
    import ctypes
    fields = [f for f, _ in ca.OzakiInfo._fields_]
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"cosma_amd.h\"\nint main(void){" + \
        'printf("%zu\\n", sizeof(cosma_amd_ozaki_info));' + "".join(
            f'printf("%zu\\n", offsetof(cosma_amd_ozaki_info, {f}));' for f in fields) + "return 0;}\n"
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "layout")
        r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-x", "c", "-", "-o", exe],
                           input=src, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        values = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert values[0] == ctypes.sizeof(ca.OzakiInfo)
    assert values[1:] == [getattr(ca.OzakiInfo, f).offset for f in fields]


# The compiler build the device code was validated with.  i8gemm.hip's BF16
# residue reduction reads and writes the accumulators in inline assembly and
# places the wait states around it by hand (the compiler's hazard recognizer
# does not look into asm blocks); its results were checked bit-identical to
# the INT8 path on the GPU (tests/test_ozaki_gpu.py, BF16 cases) with this
# compiler only.
VALIDATED_COMPILER = "roc-7.2.0 26014"


def test_device_compiler_is_the_validated_one():
    from cosma_amd import build as cbuild
    out = subprocess.run([cbuild.HIPCC, "--version"], capture_output=True, text=True).stdout
    assert VALIDATED_COMPILER in out, (
        "hipcc differs from the compiler the BF16 reduction's hand-placed wait states were validated with; "
        "re-run the BF16 split-precision GPU tests before trusting this build:\n" + out)
