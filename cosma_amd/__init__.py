"""cosma_amd — MI355X (gfx950) single-rank local-multiply path for COSMA.

Python binding of the C ABI declared in include/cosma_amd.h, loaded with
ctypes from the in-tree libcosma_amd.so.  Torch is used only to own device
memory and streams in tests and benchmarks; every compute call goes through
the HIP kernels of the library.  There is no CPU fallback: if the library is
missing, importing `lib()` raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcosma_amd.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_HIP = 2
ERR_ALLOC = 3
ERR_UNSUPPORTED = 4
ERR_NO_DEVICE = 5

_lib = None

c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_char = ctypes.c_char


class CosmaAmdError(RuntimeError):
    """Raised when a C ABI call returns a non-zero status."""

    def __init__(self, status, text):
        super().__init__(f"cosma_amd status {status}: {text}")
        self.status = status


OZAKI_MODULI = 0
OZAKI_SLICES = 1
OZAKI_MODULI_BF16 = 2
OZAKI_STRICT = 0x100   # flag: the tolerance must hold for the rigorous bound
OZAKI_SPLIT_MIN_DIM = 2048   # grouped split plans: smaller problems run in FP64 (COSMA_AMD_OZAKI_SPLIT_MIN_DIM)


class Block(ctypes.Structure):
    """One local block of a grid layout (cosma_amd_block)."""
    _fields_ = [("data", c_vp), ("ld", c_i64), ("row", c_i64), ("col", c_i64),
                ("rows", c_i64), ("cols", c_i64)]


def blocks(items):
    """ctypes array of Block from (data_ptr, ld, row, col, rows, cols) tuples."""
    arr = (Block * max(1, len(items)))()
    for i, it in enumerate(items):
        arr[i] = Block(*it)
    return arr, len(items)


class GemmDesc(ctypes.Structure):
    """One problem of a grouped GEMM (cosma_amd_gemm_desc)."""
    _fields_ = [("transa", ctypes.c_char), ("transb", ctypes.c_char), ("m", c_i64), ("n", c_i64),
                ("k", c_i64), ("alpha", c_dbl), ("a", c_vp), ("lda", c_i64), ("b", c_vp), ("ldb", c_i64),
                ("beta", c_dbl), ("c", c_vp), ("ldc", c_i64)]


class OzakiInfo(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("count", ctypes.c_int), ("bits", ctypes.c_int),
                ("int8_products", ctypes.c_int64), ("error_bound", ctypes.c_double),
                ("error_estimate", ctypes.c_double), ("tolerance_met", ctypes.c_int), ("passes", ctypes.c_int),
                ("product_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# name -> (restype, argtypes)
SIGNATURES = {
    "cosma_amd_last_error": (ctypes.c_char_p, []),
    "cosma_amd_version": (ctypes.c_char_p, []),
    "cosma_amd_handle_create": (ctypes.c_int, [ctypes.POINTER(c_vp), ctypes.c_int, ctypes.c_int, c_i64, c_i64, c_i64]),
    "cosma_amd_handle_destroy": (ctypes.c_int, [c_vp]),
    "cosma_amd_handle_get_config": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                   ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                                   ctypes.POINTER(c_i64)]),
    "cosma_amd_dgemm": (ctypes.c_int, [c_vp, c_char, c_char, c_i64, c_i64, c_i64, c_dbl,
                                       c_vp, c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp]),
    "cosma_amd_fill_uniform": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, ctypes.c_uint64, ctypes.c_uint64, c_vp]),
    "cosma_amd_dgemm_ozaki": (ctypes.c_int, [c_vp, c_char, c_char, c_i64, c_i64, c_i64, c_dbl,
                                             c_vp, c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64, ctypes.c_int,
                                             c_dbl, ctypes.c_int, ctypes.c_void_p, c_vp]),
    "cosma_amd_dgemm_ozaki_noalloc": (ctypes.c_int, [c_vp, c_char, c_char, c_i64, c_i64, c_i64, c_dbl,
                                                     c_vp, c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64, ctypes.c_int,
                                                     c_dbl, ctypes.c_int, ctypes.c_void_p, c_vp]),
    "cosma_amd_dgemm_host": (ctypes.c_int, [c_vp, c_char, c_char, c_i64, c_i64, c_i64, c_dbl,
                                            c_vp, c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64, ctypes.c_int,
                                            ctypes.c_int]),
    "cosma_amd_get_device_c": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_i64)]),
    "cosma_amd_handle_set_memory_limit": (ctypes.c_int, [c_vp, c_i64]),
    "cosma_amd_handle_last_mode": (ctypes.c_int, [c_vp]),
    "cosma_amd_fill_lognormal": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, ctypes.c_uint64, c_dbl, c_vp]),
    "cosma_amd_dgemm_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "cosma_amd_grouped_create": (ctypes.c_int, [c_vp, c_vp, c_i64, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "cosma_amd_grouped_run": (ctypes.c_int, [c_vp, c_vp]),
    "cosma_amd_grouped_graph_run": (ctypes.c_int, [c_vp, c_vp]),
    "cosma_amd_grouped_tune": (ctypes.c_int, [c_vp, c_vp]),
    "cosma_amd_grouped_set_config": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int]),
    "cosma_amd_grouped_info": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(c_i64)]),
    "cosma_amd_grouped_destroy": (ctypes.c_int, [c_vp]),
    "cosma_amd_grouped_ozaki_create": (ctypes.c_int, [c_vp, c_vp, c_i64, ctypes.c_int, ctypes.c_double,
                                                      ctypes.POINTER(c_vp)]),
    "cosma_amd_grouped_ozaki_create_ex": (ctypes.c_int, [c_vp, c_vp, c_i64, ctypes.c_int, ctypes.c_double, c_i64,
                                                         ctypes.POINTER(c_vp)]),
    "cosma_amd_grouped_ozaki_routing": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "cosma_amd_grouped_ozaki_run": (ctypes.c_int, [c_vp, c_vp]),
    "cosma_amd_grouped_ozaki_graph_run": (ctypes.c_int, [c_vp, c_vp]),
    "cosma_amd_grouped_ozaki_tune": (ctypes.c_int, [c_vp, c_vp]),
    "cosma_amd_grouped_ozaki_set_grid": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "cosma_amd_grouped_ozaki_info": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                                    ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_int)]),
    "cosma_amd_grouped_ozaki_check": (ctypes.c_int, [c_vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                                     ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_int), c_vp]),
    "cosma_amd_grouped_ozaki_destroy": (ctypes.c_int, [c_vp]),
    "cosma_amd_host_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(c_vp)]),
    "cosma_amd_test_fail_allocations": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "cosma_amd_host_free": (ctypes.c_int, [c_vp]),
    "cosma_amd_dgemm_layout": (ctypes.c_int, [c_vp, c_char, c_char, c_i64, c_i64, c_i64, c_dbl,
                                              c_vp, c_i64, c_vp, c_i64, c_dbl, c_vp, c_i64]),
    "cosma_amd_block_transform": (ctypes.c_int, [c_vp, c_char, c_i64, c_i64, c_dbl, c_vp, c_i64,
                                                 c_dbl, c_vp, c_i64, c_vp]),
}


def _load_torch_runtime_first():
    """Makes the process use one HIP runtime.

    torch ships its own libamdhip64 with the same SONAME (libamdhip64.so.7) as
    the system ROCm copy this library links against.  The dynamic loader
    serves every later request for that SONAME with whichever copy was loaded
    first, and torch only initialises on its own build, so torch (when it is
    installed) is imported before the library is loaded.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libcosma_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m cosma_amd.build` first")
        _load_torch_runtime_first()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != OK:
        raise CosmaAmdError(status, lib().cosma_amd_last_error().decode())
    return status


def _ch(t):
    return t.encode() if isinstance(t, str) else bytes([t])


class Handle:
    """Owns a cosma_amd handle (streams + workspaces) on one device."""

    def __init__(self, device=0, streams=0, tile_m=0, tile_n=0, tile_k=0):
        h = c_vp()
        check(lib().cosma_amd_handle_create(ctypes.byref(h), device, streams, tile_m, tile_n, tile_k))
        self.ptr = h

    def config(self):
        """(device, streams, tile_m, tile_n, tile_k) of the handle."""
        d, st, tm, tn, tk = ctypes.c_int(), ctypes.c_int(), c_i64(), c_i64(), c_i64()
        check(lib().cosma_amd_handle_get_config(self.ptr, ctypes.byref(d), ctypes.byref(st), ctypes.byref(tm),
                                                ctypes.byref(tn), ctypes.byref(tk)))
        return d.value, st.value, tm.value, tn.value, tk.value

    def close(self):
        if self.ptr:
            lib().cosma_amd_handle_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # Raw pointer calls (device pointers as ints)
    def dgemm(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
        return check(lib().cosma_amd_dgemm(self.ptr, _ch(transa), _ch(transb), m, n, k, alpha,
                                           a, lda, b, ldb, beta, c, ldc, stream))

    def dgemm_ozaki(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                    method=OZAKI_MODULI, tolerance=1e-14, count=0, stream=None):
        """Split-precision GEMM on device pointers; returns the OzakiInfo as a dict."""
        info = OzakiInfo()
        check(lib().cosma_amd_dgemm_ozaki(self.ptr, _ch(transa), _ch(transb), m, n, k, alpha, a, lda, b, ldb,
                                          beta, c, ldc, method, tolerance, count, ctypes.addressof(info),
                                          stream))
        return info.as_dict()

    def dgemm_ozaki_noalloc(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                            method=OZAKI_MODULI, tolerance=1e-14, count=0, stream=None):
        """The split-precision GEMM without allocation (library scratch); returns the info dict."""
        info = OzakiInfo()
        check(lib().cosma_amd_dgemm_ozaki_noalloc(self.ptr, _ch(transa), _ch(transb), m, n, k, alpha, a, lda, b,
                                                  ldb, beta, c, ldc, method, tolerance, count,
                                                  ctypes.addressof(info), stream))
        return info.as_dict()

    def dgemm_host(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                   pin_host_buffers=True, copy_c_back=True):
        """Host-streaming GEMM; a, b, c are raw pointers (host or device)."""
        return check(lib().cosma_amd_dgemm_host(self.ptr, _ch(transa), _ch(transb), m, n, k, alpha,
                                                a, lda, b, ldb, beta, c, ldc, int(pin_host_buffers),
                                                int(copy_c_back)))

    def device_c(self):
        p, ld = c_vp(), c_i64()
        check(lib().cosma_amd_get_device_c(self.ptr, ctypes.byref(p), ctypes.byref(ld)))
        return p.value, ld.value

    def set_memory_limit(self, nbytes):
        check(lib().cosma_amd_handle_set_memory_limit(self.ptr, nbytes))

    def last_mode(self):
        return lib().cosma_amd_handle_last_mode(self.ptr)

    def dgemm_layout(self, transa, transb, m, n, k, alpha, a_blocks, b_blocks, beta, c_blocks):
        """Grid-layout multiply; *_blocks are lists of (ptr, ld, row, col, rows, cols)."""
        A, na = blocks(a_blocks)
        B, nb = blocks(b_blocks)
        C, nc = blocks(c_blocks)
        check(lib().cosma_amd_dgemm_layout(self.ptr, _ch(transa), _ch(transb), m, n, k, alpha,
                                           ctypes.cast(A, c_vp), na, ctypes.cast(B, c_vp), nb, beta,
                                           ctypes.cast(C, c_vp), nc))

    def block_transform(self, op, rows, cols, alpha, src, ld_src, beta, dst, ld_dst, stream=None):
        return check(lib().cosma_amd_block_transform(self.ptr, _ch(op), rows, cols, alpha, src, ld_src,
                                                     beta, dst, ld_dst, stream))


class Grouped:
    """A grouped-GEMM plan over device-resident problems.

    problems: iterable of (transa, transb, m, n, k, alpha, a_ptr, lda, b_ptr, ldb, beta, c_ptr, ldc).
    tile_config: 0 = tuned choice for these shapes if known in the process, else 1; 1 = 64x64,
    2 = 128x64, 3 = 128x128 (LDS-staged, four waves); 4 = 64x64, 5 = 32x32 (one wave per tile,
    fragments straight from global memory).  tune() times them all with several grids.
    """

    def __init__(self, handle, problems, tile_config=0):
        probs = list(problems)
        arr = (GemmDesc * max(1, len(probs)))()
        for i, (ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) in enumerate(probs):
            arr[i] = GemmDesc(_ch(ta), _ch(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        p = c_vp()
        check(lib().cosma_amd_grouped_create(handle.ptr, ctypes.cast(arr, c_vp), len(probs), tile_config,
                                             ctypes.byref(p)))
        self.ptr = p

    def run(self, stream=None):
        check(lib().cosma_amd_grouped_run(self.ptr, stream))

    def graph_run(self, stream=None):
        check(lib().cosma_amd_grouped_graph_run(self.ptr, stream))

    def tune(self, stream=None):
        check(lib().cosma_amd_grouped_tune(self.ptr, stream))

    def set_config(self, tile_config, grid=0):
        check(lib().cosma_amd_grouped_set_config(self.ptr, tile_config, grid))

    def info(self):
        cfg, grid, tiles = ctypes.c_int(), ctypes.c_int(), c_i64()
        check(lib().cosma_amd_grouped_info(self.ptr, ctypes.byref(cfg), ctypes.byref(grid), ctypes.byref(tiles)))
        return {"tile_config": cfg.value, "grid": grid.value, "tiles": tiles.value}

    def close(self):
        if self.ptr:
            lib().cosma_amd_grouped_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GroupedOzaki:
    """A grouped split-precision (moduli method) plan over device-resident problems.

    problems: iterable of (transa, transb, m, n, k, alpha, a_ptr, lda, b_ptr, ldb, beta, c_ptr, ldc);
    moduli = 0 chooses the count from `tolerance` on the operands present now.  Problems with
    min(m, n, k) < split_min_dim are computed by the FP64 grouped kernel in the same run (None:
    the library default, SPLIT_MIN_DIM; 0: every problem is split).
    """

    def __init__(self, handle, problems, moduli=0, tolerance=1e-14, split_min_dim=None):
        probs = list(problems)
        arr = (GemmDesc * max(1, len(probs)))()
        for i, (ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) in enumerate(probs):
            arr[i] = GemmDesc(_ch(ta), _ch(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        p = c_vp()
        split = OZAKI_SPLIT_MIN_DIM if split_min_dim is None else split_min_dim
        check(lib().cosma_amd_grouped_ozaki_create_ex(handle.ptr, ctypes.cast(arr, c_vp), len(probs), moduli,
                                                      tolerance, split, ctypes.byref(p)))
        self.ptr = p

    def run(self, stream=None):
        check(lib().cosma_amd_grouped_ozaki_run(self.ptr, stream))

    def graph_run(self, stream=None):
        check(lib().cosma_amd_grouped_ozaki_graph_run(self.ptr, stream))

    def tune(self, stream=None):
        check(lib().cosma_amd_grouped_ozaki_tune(self.ptr, stream))

    def set_grid(self, grid):
        check(lib().cosma_amd_grouped_ozaki_set_grid(self.ptr, grid))

    def info(self):
        mod, est, tiles, grid = ctypes.c_int(), ctypes.c_double(), c_i64(), ctypes.c_int()
        check(lib().cosma_amd_grouped_ozaki_info(self.ptr, ctypes.byref(mod), ctypes.byref(est), ctypes.byref(tiles),
                                                 ctypes.byref(grid)))
        sp, fp = c_i64(), c_i64()
        check(lib().cosma_amd_grouped_ozaki_routing(self.ptr, ctypes.byref(sp), ctypes.byref(fp)))
        return {"moduli": mod.value, "error_estimate": est.value, "tiles": tiles.value, "grid": grid.value,
                "split_problems": sp.value, "fp64_problems": fp.value}

    def check(self, tolerance=1e-14, stream=None):
        """Error check of the last run: worst expected error, its problem, moduli needed."""
        w, i, need = ctypes.c_double(), c_i64(), ctypes.c_int()
        check(lib().cosma_amd_grouped_ozaki_check(self.ptr, tolerance, ctypes.byref(w), ctypes.byref(i),
                                                  ctypes.byref(need), stream))
        return {"worst_estimate": w.value, "worst_problem": i.value, "moduli_needed": need.value}

    def close(self):
        if self.ptr:
            lib().cosma_amd_grouped_ozaki_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_uniform(ptr, rows, cols, ld, seed, offset=0, stream=None):
    """Seeded uniform [-1, 1) fill of a column-major device matrix."""
    check(lib().cosma_amd_fill_uniform(ptr, rows, cols, ld, seed, offset, stream))


def fill_lognormal(ptr, rows, cols, ld, seed, phi, stream=None):
    """a = (u - 0.5) * exp(phi * z) fill of a column-major device matrix."""
    check(lib().cosma_amd_fill_lognormal(ptr, rows, cols, ld, seed, phi, stream))


def fail_allocations(skip, count):
    """Lets the library's next `skip` device or page-locked allocations through and fails the
    `count` after them (count 0 disarms); for tests of the error paths."""
    check(lib().cosma_amd_test_fail_allocations(skip, count))


def set_dgemm_tuning(tile_config=0, group_m=0):
    """Selects the DGEMM tile shape (0 = automatic)."""
    check(lib().cosma_amd_dgemm_set_tuning(tile_config, group_m))


def colmajor(t):
    """Returns (ptr, ld) of a torch tensor viewed as a column-major matrix.

    A column-major r x c matrix is stored here as a torch tensor of shape
    (c, ld) (row i of the tensor = column i of the matrix), so that
    element (row, col) lives at ptr[row + col*ld].
    """
    assert t.dim() == 2 and t.stride(1) == 1
    return t.data_ptr(), t.stride(0)
