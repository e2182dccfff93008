"""Benchmark of the single-rank local multiply on one MI355X.

Headline (`value`): C <- alpha * A^T * B + beta * C with m = n = k = 16384,
FP64, A, B, C uniform [-1, 1) (seeded SplitMix64, generated in HBM),
alpha = 1, beta = 0.5, operands device-resident.  One step = one DGEMM call
through the C ABI; the line carries the kernel's roofline fraction (HIP
events on its launch stream), the error of the result against the
double-double oracle on sampled rows, and the reference's CPU path (one
host-BLAS cblas_dgemm, src/cosma/blas.cpp:35) timed on this machine's cores.

The same JSON line carries every other benchmark configuration as a
sub-object with its own time, error and roofline (`configs`):
  1  2000^3 NN through cosma::multiply with CosmaMatrix objects (the
     reference's miniapp shape, cosma_amd/bin/cosma_miniapp) beside the same
     product by host BLAS;
  2  the headline GEMM from pinned host memory in 4096^2 tiles, and the
     layout transforms of a 16384^2 pdgemm on the device;
  3  tall-skinny A^T B, m = n = 2048, k = 2^20, streamed from pinned host
     memory in 65536-deep k-panels;
  4  8192^3 split precision on a = (u - 1/2) exp(phi z), phi in {0, 1, 4},
     INT8 residues, BF16 residues and slices, against the FP64 DGEMM;
  5  4096 uneven problems (m, n, k log-uniform in [16, 512]) in one grouped
     launch replayed as a hipGraph, plain and split precision.
`--quick` keeps only the headline, the split-precision leg at the headline
shape and the CPU baseline; `--workload X` runs one leg alone.
"""
import argparse
import gc
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6   # TFLOP/s, MI355X datasheet dense FP64 matrix
PEAK_INT8 = 5000.0      # TOP/s, MI355X datasheet dense INT8 matrix
PEAK_BF16 = 2500.0      # TFLOP/s, MI355X datasheet dense BF16 matrix
PEAK_HBM = 8000.0       # GB/s, MI355X datasheet
M = N = K = 16384
ALPHA, BETA = 1.0, 0.5
METRIC = "FP64-accurate TFLOP/s, 1 GPU: native DGEMM vs split-precision at stated error"
LEGS = ["device", "host", "tall-skinny", "ozaki", "ozaki-8192", "ozaki-target", "reorder", "batched", "miniapp",
        "vendor"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default=None, choices=LEGS,
                   help="run one leg alone and print its line: device (the headline with the split-precision "
                        "leg and the CPU baseline), host (16384^3 from pinned host memory), tall-skinny, ozaki "
                        "(8192^3 phi sweep with the error of every count), ozaki-8192 (the 8192^3 moduli call "
                        "alone at --phi, for profiling), ozaki-target (split precision at the "
                        "headline shape), reorder, batched, miniapp (2000^3 through cosma::multiply), vendor "
                        "(torch's own FP64, INT8 and BF16 GEMMs on this GPU, as reference points)")
    p.add_argument("--phi", type=float, default=0.0, help="phi of the ozaki-8192 leg's operands")
    p.add_argument("--quick", action="store_true", help="headline, split precision and CPU baseline only")
    p.add_argument("--no-split", action="store_true", help="skip the split-precision summary leg")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    p.add_argument("--cpu-n", type=int, default=16384, help="CPU baseline size (m = n = k)")
    p.add_argument("--no-check", action="store_true", help="skip the sampled-row error checks")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a 2 GiB device copy (block_transform 'N', 2 GiB read + 2 GiB "
                        "written) to calibrate HBM byte counters")
    return p.parse_args()


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profile():
    """HBM bytes per launch of the DGEMM kernel from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "dgemm_tn_16384_pmc.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def host_threads():
    """Host threads for the CPU legs: the machine's share for one GPU (16),
    or fewer if the process may use fewer cores."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(n):
    """The reference CPU path (one host-BLAS dgemm, op(A) = A^T, beta = 0.5) at
    the headline shape: one warm-up at n / 2, then two timed calls at n."""
    import oracle
    threads = host_threads()
    oracle.hostblas.set_threads(threads)
    say(f"cpu baseline: cblas_dgemm TN {n}^3 on {threads} threads")
    w = n // 2
    A, B, C = oracle.uniform(w, w, 1), oracle.uniform(w, w, 2), oracle.uniform(w, w, 3)
    oracle.hostblas.dgemm("T", "N", w, w, w, ALPHA, A, w, B, w, BETA, C, w)   # warm-up (threads, pages)
    del A, B, C
    A, B, C = oracle.uniform(n, n, 1), oracle.uniform(n, n, 2), oracle.uniform(n, n, 3)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        oracle.hostblas.dgemm("T", "N", n, n, n, ALPHA, A, n, B, n, BETA, C, n)
        times.append(time.perf_counter() - t0)
    best = min(times)
    return {"value": round(2.0 * n ** 3 / best / 1e12, 4), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)),
            "sample": f"cblas_dgemm('T','N') m=n=k={n} (the headline shape), alpha=1, beta=0.5, uniform[-1,1) "
                      f"inputs, min of {len(times)} calls after a warm-up at {w}^3, BLAS "
                      f"{oracle.hostblas.vendor()} on {threads} threads (the host share of one GPU on this "
                      f"machine; OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')})"}


def sampled_error(A, B, C, C0_seed, rows=(0, 8191)):
    """Relative error of sampled rows of C (headline product) against the double-double oracle."""
    import numpy as np
    import oracle
    from oracle.sampled import dd_lines, lines_error
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()           # tensors (m, k) / (n, k): A stored k x m
    r, _ = dd_lines("T", "N", M, N, K, Ah.ctypes.data, K, Bh.ctypes.data, K, rows, ())
    got, ref = {}, {}
    for i in rows:
        c0 = oracle.uniform_at(C0_seed, i + np.arange(N, dtype=np.uint64) * M)
        ref[i] = ALPHA * r[i] + BETA * c0
        got[i] = C[:, i].cpu().numpy()
    return lines_error(got, {}, ref, {})


def link_bandwidth(torch, nbytes=1 << 31):
    """Pinned host -> device copy rate in GB/s (the host-link roofline)."""
    src = torch.empty(nbytes // 8, dtype=torch.float64, pin_memory=True)
    dst = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    return 3 * nbytes / (time.perf_counter() - t0) / 1e9


def fill_pinned(t, seed):
    """Uniform fill of a pinned host tensor (cols, rows), generated on the GPU in pieces."""
    import torch
    import cosma_amd as ca
    cols, rows = t.shape
    step = max(1, (1 << 28) // rows)
    tmp = torch.empty((min(step, cols), rows), dtype=torch.float64, device="cuda")
    for c0 in range(0, cols, step):
        cc = min(step, cols - c0)
        ca.fill_uniform(tmp.data_ptr(), rows, cc, rows, seed, offset=c0 * rows)
        t[c0:c0 + cc].copy_(tmp[:cc])
    torch.cuda.synchronize()
    del tmp


def streamed_error(h, ta, tb, m, n, k, alpha, beta, host, C0_seed):
    """One more call from the seeded C0, then the relative error of rows 0 and
    m-1 and column n/2 of C against the double-double oracle (all k terms)."""
    from oracle.sampled import dd_lines, lines_error
    C = host["C"]                                   # torch (n, m): row j = column j of C
    fill_pinned(C, C0_seed)
    rows, cols = (0, m - 1), (n // 2,)
    c0r = {i: C[:, i].numpy().copy() for i in rows}
    c0c = {j: C[j, :].numpy().copy() for j in cols}
    h.dgemm_host(ta, tb, m, n, k, alpha, host["A"].data_ptr(), k, host["B"].data_ptr(), k, beta, C.data_ptr(), m)
    r, c = dd_lines(ta, tb, m, n, k, host["A"].data_ptr(), k, host["B"].data_ptr(), k, rows, cols)
    return lines_error({i: C[:, i].numpy() for i in rows}, {j: C[j, :].numpy() for j in cols},
                       {i: alpha * r[i] + beta * c0r[i] for i in rows},
                       {j: alpha * c[j] + beta * c0c[j] for j in cols})


def streamed(kind, steps, warmup, check=True):
    """Host-resident workloads through the host-streaming DGEMM."""
    import torch
    import cosma_amd as ca
    if kind == "host":
        m = n = k = 16384
        ta, tb, alpha, beta = "T", "N", 1.0, 0.5
        tiles = (4096, 4096, 4096)
        desc = "dgemm op(A)=A^T m=n=k=16384 alpha=1 beta=0.5, pinned host operands, 4096^2 tiles"
    else:
        m = n = 2048
        k = 1 << 20
        ta, tb, alpha, beta = "T", "N", 1.0, 1.0
        tiles = (4096, 4096, 65536)
        desc = "dgemm op(A)=A^T m=n=2048 k=1048576 alpha=beta=1, pinned host operands, k-panels of 65536"
    say(f"streamed leg: {desc}")
    h = ca.Handle(0, streams=2, tile_m=tiles[0], tile_n=tiles[1], tile_k=tiles[2])
    shapes = {"A": (m, k), "B": (n, k), "C": (n, m)}            # torch (cols, rows); A stored k x m
    host = {x: torch.empty(sh, dtype=torch.float64, pin_memory=True) for x, sh in shapes.items()}
    for seed, x in enumerate(shapes, start=11):
        fill_pinned(host[x], seed)
    C0_seed = 13
    link = link_bandwidth(torch)

    def call():
        h.dgemm_host(ta, tb, m, n, k, alpha, host["A"].data_ptr(), k, host["B"].data_ptr(), k, beta,
                     host["C"].data_ptr(), m)

    times = []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        call()
        dt = time.perf_counter() - t0
        if i >= warmup:
            times.append(dt)
    t = sum(times) / len(times)
    mode = h.last_mode()
    err = streamed_error(h, ta, tb, m, n, k, alpha, beta, host, C0_seed) if check else None
    flops = 2.0 * m * n * k
    h2d = 8.0 * (m * k + k * n + (m * n if beta != 0 else 0))
    d2h = 8.0 * m * n
    t_mfma = flops / (PEAK_FP64_MFMA * 1e12)
    t_link = h2d / (link * 1e9)
    out = {
        "value": round(flops / t / 1e12, 3), "unit": "TFLOP/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(t * 1e3, 3),
        "config": {"workload": desc, "m": m, "n": n, "k": k, "tiles": tiles, "schedule":
                   {0: "device", 1: "resident panels", 2: "bounded ring"}[mode]},
        "error": {"rel_fro_sampled_lines_vs_double_double": err, "bound_4k_u": 4 * k * 2.0 ** -53,
                  "lines": "rows 0, m-1 and column n/2 of C, each with all k terms"},
        "roofline": {"bound": "host-link" if t_link > t_mfma else "mfma",
                     "achieved": round(h2d / t / 1e9, 2), "peak": round(link, 2), "unit": "GB/s",
                     "frac": round(max(t_link, t_mfma) / t, 4), "traffic": h2d + d2h,
                     "note": "achieved = host->device bytes / time; peak = measured pinned H2D copy rate; "
                             "frac = max(link time, MFMA time) / time"},
    }
    h.close()
    del host
    gc.collect()
    torch.cuda.empty_cache()
    return out


def dd_rows_nn(Ad, Bd, n, rows):
    """Rows of A*B (NN, n x n column-major device operands) from the double-double oracle."""
    from oracle.sampled import dd_lines
    Ah, Bh = Ad.cpu().numpy(), Bd.cpu().numpy()
    r, _ = dd_lines("N", "N", n, n, n, Ah.ctypes.data, n, Bh.ctypes.data, n, rows, ())
    return [r[i] for i in rows]


def dd_rows_tn(Ad, Bd, n, rows):
    """Rows of A^T*B (A stored k x m, B k x n, column-major device operands)."""
    from oracle.sampled import dd_lines
    Ah, Bh = Ad.cpu().numpy(), Bd.cpu().numpy()
    r, _ = dd_lines("T", "N", n, n, n, Ah.ctypes.data, n, Bh.ctypes.data, n, rows, ())
    return [r[i] for i in rows]


def rows_error(Cd, rows, ref_rows):
    """Relative Frobenius error of sampled rows of C (device, column-major)."""
    num = den = 0.0
    for i, ref in zip(rows, ref_rows):
        got = Cd[:, i].cpu().numpy()
        num += float(((got - ref) ** 2).sum())
        den += float((ref ** 2).sum())
    return (num / den) ** 0.5


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def full_diff(C, C_fp64):
    """Relative Frobenius difference of the whole matrix C against the FP64
    GPU result (itself parity-tested against the oracle), on the device: it
    bounds every element, where the oracle comparison samples rows."""
    import torch
    d = torch.linalg.vector_norm(C - C_fp64).item()
    r = torch.linalg.vector_norm(C_fp64).item()
    ok = bool(torch.isfinite(C).all().item())
    return {"rel_fro_whole_matrix_vs_fp64_gpu": d / r if r > 0 else d, "all_finite": ok}


def ozaki_entry(info, t, t_plain, flops, err, n):
    """Per-method record of a split-precision leg, with the product kernel's
    roofline (INT8 or BF16 matrix peak; 2 n^3 ops per product)."""
    bf16 = info["method"] == 2
    ops = 2.0 * n ** 3 * info["int8_products"]
    prod = ops / (info["product_ms"] * 1e-3) / 1e12 if info["product_ms"] > 0 else None
    peak = PEAK_BF16 if bf16 else PEAK_INT8
    return {"tflops_fp64_equiv": round(flops / t / 1e12, 3), "ms": round(t * 1e3, 3),
            "speedup_vs_fp64": round(t_plain / t, 3), "rel_err": err, "count": info["count"], "bits": info["bits"],
            "products": info["int8_products"], "error_estimate": info["error_estimate"],
            "error_bound": info["error_bound"], "tolerance_met": info["tolerance_met"], "passes": info["passes"],
            "roofline": {"bound": "mfma", "kernel": "i8gemm_kernel (all products of the call)",
                         "achieved": round(prod, 1) if prod else None, "peak": peak,
                         "unit": "TFLOP/s" if bf16 else "TOP/s", "frac": round(prod / peak, 4) if prod else None,
                         "traffic": None}}


def split_precision(phis=(0.0, 1.0, 4.0), steps=5, warmup=1, n=8192, rows=(0, 4097), sweep=True,
                    methods=("moduli", "moduli_bf16", "slices")):
    """8192^3 C = A*B (NN): plain FP64 MFMA DGEMM vs the split-precision DGEMM
    (moduli at the default 1e-14 tolerance with INT8 and with BF16 residues,
    slices at the default tolerance), inputs a = (u - 0.5) exp(phi z).  Errors
    are relative Frobenius errors of sampled rows against the double-double
    oracle, with the whole-matrix difference against the FP64 GPU result.
    With `sweep`, also the error and time of every count."""
    import torch
    import cosma_amd as ca
    h = ca.Handle(0)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    B = torch.empty((n, n), dtype=torch.float64, device="cuda")
    C = torch.empty((n, n), dtype=torch.float64, device="cuda")
    flops = 2.0 * n ** 3
    results = []
    for phi in phis:
        say(f"split precision 8192^3 phi={phi}")
        ca.fill_lognormal(A.data_ptr(), n, n, n, 21, phi)
        ca.fill_lognormal(B.data_ptr(), n, n, n, 22, phi)
        torch.cuda.synchronize()
        ref = dd_rows_nn(A, B, n, rows)
        plain = lambda: h.dgemm("N", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n)
        t_plain = timed(plain, steps, warmup)
        entry = {"phi": phi, "fp64": {"tflops": round(flops / t_plain / 1e12, 3), "ms": round(t_plain * 1e3, 3),
                                      "rel_err": rows_error(C, rows, ref)}}
        C_fp64 = C.clone()
        for name, method in (("moduli", ca.OZAKI_MODULI), ("moduli_bf16", ca.OZAKI_MODULI_BF16),
                             ("slices", ca.OZAKI_SLICES)):
            if name not in methods:
                continue
            info = {}

            def oz():
                info.update(h.dgemm_ozaki("N", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0,
                                          C.data_ptr(), n, method=method, tolerance=1e-14))
            t = timed(oz, steps, warmup)
            entry[name] = ozaki_entry(info, t, t_plain, flops, rows_error(C, rows, ref), n)
            entry[name].update(full_diff(C, C_fp64))
        del C_fp64
        if sweep:
            # error against the double-double oracle for each count: moduli 8..20
            # and slices 4..8 (one warm-up and one timed call each)
            sw = []
            for name, method, counts in (("moduli", ca.OZAKI_MODULI, range(8, 21, 2)),
                                         ("slices", ca.OZAKI_SLICES, range(4, 9))):
                for cnt in counts:
                    info = {}

                    def oz():
                        info.update(h.dgemm_ozaki("N", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0,
                                                  C.data_ptr(), n, method=method, count=cnt))
                    t = timed(oz, 1, 1)
                    sw.append({"method": name, "count": cnt, "bits": info["bits"], "ms": round(t * 1e3, 3),
                               "rel_err": rows_error(C, rows, ref), "error_estimate": info["error_estimate"],
                               "error_bound": info["error_bound"]})
            entry["count_sweep"] = sw
        results.append(entry)
    h.close()
    del A, B, C
    torch.cuda.empty_cache()
    return results


def split_precision_target(steps=3, warmup=1, n=16384, rows=(0, 8193)):
    """Split precision at the plain FP64 shape: m = n = k = 16384, op(A) = A^T,
    alpha = 1, uniform [-1, 1) inputs, beta = 0 (so that C needs no restoring
    between timed calls): plain FP64 MFMA DGEMM vs the moduli method at the
    default 1e-14 tolerance, errors of sampled rows against the double-double
    oracle and the whole-matrix difference against the FP64 GPU result."""
    import torch
    import cosma_amd as ca
    say("split precision at the headline shape")
    h = ca.Handle(0)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    B = torch.empty((n, n), dtype=torch.float64, device="cuda")
    C = torch.empty((n, n), dtype=torch.float64, device="cuda")
    ca.fill_uniform(A.data_ptr(), n, n, n, 41)
    ca.fill_uniform(B.data_ptr(), n, n, n, 42)
    torch.cuda.synchronize()
    ref = dd_rows_tn(A, B, n, rows)
    flops = 2.0 * n ** 3
    t_plain = timed(lambda: h.dgemm("T", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n),
                    steps, warmup)
    e_plain = rows_error(C, rows, ref)
    C_fp64 = C.clone()
    info = {}

    def oz():
        info.update(h.dgemm_ozaki("T", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n,
                                  method=ca.OZAKI_MODULI, tolerance=1e-14))
    t_oz = timed(oz, steps, warmup)
    out = {"workload": "C=A^T*B m=n=k=16384 alpha=1 beta=0, uniform[-1,1) inputs",
           "fp64": {"tflops": round(flops / t_plain / 1e12, 3), "ms": round(t_plain * 1e3, 3), "rel_err": e_plain},
           "moduli": ozaki_entry(info, t_oz, t_plain, flops, rows_error(C, rows, ref), n)}
    out["moduli"].update(full_diff(C, C_fp64))
    h.close()
    del A, B, C, C_fp64
    torch.cuda.empty_cache()
    return out


def reorder(steps, warmup):
    """Device block transforms of the ScaLAPACK -> COSMA -> ScaLAPACK round trip
    of a 16384^2 pdgemm on one rank: A transposed (op = 'T'), B copied
    (op = 'N'), the result written back as C <- alpha * X + beta * C; the
    transposed copy is checked bit for bit on sampled columns."""
    import numpy as np
    import torch
    import cosma_amd as ca
    say("reorder leg: 16384^2 transforms")
    n = 16384
    h = ca.Handle(0)
    X = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Y = torch.empty((n, n), dtype=torch.float64, device="cuda")
    Z = torch.empty((n, n), dtype=torch.float64, device="cuda")
    ca.fill_uniform(X.data_ptr(), n, n, n, 31)
    ca.fill_uniform(Y.data_ptr(), n, n, n, 32)
    ca.fill_uniform(Z.data_ptr(), n, n, n, 33)
    torch.cuda.synchronize()
    nb = 8.0 * n * n
    legs = {"A transpose (alpha=1, beta=0)": (lambda: h.block_transform("T", n, n, 1.0, X.data_ptr(), n, 0.0,
                                                                          Y.data_ptr(), n), 2 * nb),
            "B copy (alpha=1, beta=0)": (lambda: h.block_transform("N", n, n, 1.0, X.data_ptr(), n, 0.0,
                                                                     Y.data_ptr(), n), 2 * nb),
            "C epilogue (alpha=1, beta=0.5)": (lambda: h.block_transform("N", n, n, 1.0, X.data_ptr(), n, 0.5,
                                                                           Z.data_ptr(), n), 3 * nb)}
    res, total_b, total_t = {}, 0.0, 0.0
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, (fn, nbytes) in legs.items():
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        s.record()
        for _ in range(steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / steps
        res[name] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                     "frac_of_hbm": round(nbytes / ms / 1e6 / PEAK_HBM, 4)}
        total_b += nbytes
        total_t += ms
    # bit-exactness of the transpose on sampled columns of the result
    h.block_transform("T", n, n, 1.0, X.data_ptr(), n, 0.0, Y.data_ptr(), n)
    torch.cuda.synchronize()
    exact = all(np.array_equal(Y[j].cpu().numpy(), X[:, j].cpu().numpy()) for j in (0, 4097, n - 1))
    h.close()
    del X, Y, Z
    torch.cuda.empty_cache()
    return {"value": round(total_b / total_t / 1e6, 1), "unit": "GB/s", "steps": steps, "warmup": warmup,
            "ms_per_step": round(total_t, 4),
            "config": {"workload": "16384^2 ScaLAPACK <-> COSMA transforms of one pdgemm on one rank"},
            "legs": res, "bit_exact_transpose_sampled_columns": exact,
            "roofline": {"bound": "hbm", "achieved": round(total_b / total_t / 1e6, 1), "peak": PEAK_HBM,
                         "unit": "GB/s", "frac": round(total_b / total_t / 1e6 / PEAK_HBM, 4), "traffic": None}}


def batch_dims(count=4096, seed=5):
    """m, n, k of the batched workload: each round(exp(U[ln 16, ln 512]))
    clamped to [16, 512], numpy default_rng(seed); shape (count, 3)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(np.exp(rng.uniform(np.log(16), np.log(512), (count, 3)))), 16, 512).astype(np.int64)


def batched(steps, warmup, count=4096, seed=5):
    """4096 independent C = A B problems (NN, alpha = 1, beta = 0) packed in
    three device buffers; one grouped persistent launch per step, replayed as
    a hipGraph after the plan is tuned; then the same batch in split-precision
    form (grouped moduli plan at the 1e-14 tolerance) with its check, with
    the default split threshold (every problem here runs in FP64) and with
    every problem split."""
    import numpy as np
    import torch
    import cosma_amd as ca
    import oracle
    say("batched leg: 4096 problems")
    dims = batch_dims(count, seed)

    def offsets(sizes):
        sizes = (sizes + 1) // 2 * 2                      # keep every operand 16-byte aligned
        return np.concatenate([[0], np.cumsum(sizes)])
    m, n, k = dims[:, 0], dims[:, 1], dims[:, 2]
    oa, ob, oc = offsets(m * k), offsets(k * n), offsets(m * n)
    A = torch.empty(int(oa[-1]), dtype=torch.float64, device="cuda")
    B = torch.empty(int(ob[-1]), dtype=torch.float64, device="cuda")
    C = torch.empty(int(oc[-1]), dtype=torch.float64, device="cuda")
    ca.fill_uniform(A.data_ptr(), int(oa[-1]), 1, int(oa[-1]), 41)
    ca.fill_uniform(B.data_ptr(), int(ob[-1]), 1, int(ob[-1]), 42)
    torch.cuda.synchronize()
    pa, pb, pc = A.data_ptr(), B.data_ptr(), C.data_ptr()
    probs = [("N", "N", int(m[i]), int(n[i]), int(k[i]), 1.0, pa + 8 * int(oa[i]), int(m[i]), pb + 8 * int(ob[i]),
              int(k[i]), 0.0, pc + 8 * int(oc[i]), int(m[i])) for i in range(count)]
    h = ca.Handle(0)
    plan = ca.Grouped(h, probs)
    plan.tune()
    for _ in range(warmup):
        plan.graph_run()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(steps):
        plan.graph_run()
    e.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ms = s.elapsed_time(e) / steps
    flops = 2.0 * (m * n * k).astype(np.float64)
    nbytes = 8.0 * (m * k + k * n + m * n).astype(np.float64)
    t_roof = float(np.maximum(flops / (PEAK_FP64_MFMA * 1e12), nbytes / (PEAK_HBM * 1e9)).sum())
    # correctness of a sample of problems against the double-double oracle
    sample = sorted(set([0, 1, count // 2, count - 1] + np.argsort(m * n * k)[-4:].tolist()))
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    refs = {}
    for i in sample:
        ref = np.zeros((int(m[i]), int(n[i])), order="F")
        oracle.dgemm_raw("N", "N", int(m[i]), int(n[i]), int(k[i]), 1.0, Ah.ctypes.data + 8 * int(oa[i]), int(m[i]),
                         Bh.ctypes.data + 8 * int(ob[i]), int(k[i]), 0.0, ref.ctypes.data, int(m[i]), dd=True)
        refs[i] = ref

    def sampled_err():
        Ch = C.cpu().numpy()
        return max(oracle.rel_fro(Ch[int(oc[i]):int(oc[i]) + int(m[i] * n[i])].reshape((int(n[i]), int(m[i]))).T,
                                  refs[i]) for i in sample)

    err = sampled_err()
    info = plan.info()
    value = float(flops.sum()) / wall / 1e12
    out = {"value": round(value, 3), "unit": "TFLOP/s", "steps": steps, "warmup": warmup,
           "ms_per_step": round(wall * 1e3, 4),
           "config": {"workload": f"{count} NN problems, m,n,k log-uniform in [16,512], one grouped launch per "
                                  f"step replayed as a hipGraph", "sum_flops": float(flops.sum()),
                      "sum_bytes": float(nbytes.sum()), "plan": info},
           "error": {"max_rel_fro_sampled_problems_vs_double_double": err, "sampled_problems": len(sample)},
           "roofline": {"bound": "per-problem max(MFMA, HBM)",
                        "achieved": round(float(flops.sum()) / (ms * 1e-3) / 1e12, 3),
                        "peak": round(float(flops.sum()) / t_roof / 1e12, 3), "unit": "TFLOP/s",
                        "frac": round(t_roof / (ms * 1e-3), 4), "traffic": None,
                        "note": "peak = sum flops / sum_i max(flops_i / 78.6 TF, bytes_i / 8 TB/s); "
                                "frac uses the launch-inclusive event time of a graph replay"}}
    plan.close()
    # the split-precision form: problems below the split threshold (all of
    # this batch) run on the FP64 grouped kernel; then every problem split
    for key, split_min in (("ozaki_form", None), ("ozaki_form_split_all", 0)):
        oz = ca.GroupedOzaki(h, probs, split_min_dim=split_min)
        oz.tune()
        for _ in range(warmup):
            oz.graph_run()
        torch.cuda.synchronize()
        s.record()
        for _ in range(steps):
            oz.graph_run()
        e.record()
        torch.cuda.synchronize()
        oz_ms = s.elapsed_time(e) / steps
        out[key] = {"ms_per_step": round(oz_ms, 4),
                    "tflops_fp64_equiv": round(float(flops.sum()) / (oz_ms * 1e-3) / 1e12, 3),
                    "vs_plain_fp64": round(ms / oz_ms, 3), "plan": oz.info(), "check": oz.check(1e-14),
                    "max_rel_fro_sampled_problems_vs_double_double": sampled_err()}
        oz.close()
    h.close()
    del A, B, C
    torch.cuda.empty_cache()
    return out


def miniapp(reps=5, n=2000):
    """The reference's miniapp shape: C = A B, m = n = k = 2000, NN, alpha = 1,
    beta = 0, through cosma::multiply with CosmaMatrix objects on one rank
    (cosma_amd/bin/cosma_miniapp: host matrices in the page-locked pool,
    streamed to the GPU), and the same product by host BLAS (the reference's
    CPU path), with the error of sampled lines of C."""
    import numpy as np
    import oracle
    from oracle.sampled import dd_lines, lines_error
    from cosma_amd import build as cbuild
    say("miniapp leg: 2000^3 through cosma::multiply")
    exe = cbuild.MINIAPP
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built")
    with tempfile.TemporaryDirectory() as d:
        out_path = os.path.join(d, "c.bin")
        r = subprocess.run([exe, "-m", str(n), "-n", str(n), "-k", str(n), "-r", str(reps), "--json", "--out",
                            out_path], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"cosma_miniapp failed: {r.stderr}")
        gpu = json.loads(r.stdout.strip().splitlines()[-1])
        Cg = np.fromfile(out_path, dtype=np.float64).reshape((n, n), order="F")
    A, B = oracle.uniform(n, n, 1), oracle.uniform(n, n, 2)
    rows, cols = (0, n - 1), (n // 2,)
    rr, cc = dd_lines("N", "N", n, n, n, A.ctypes.data, n, B.ctypes.data, n, rows, cols)
    err = lines_error({i: Cg[i, :] for i in rows}, {j: Cg[:, j] for j in cols}, rr, cc)
    threads = host_threads()
    oracle.hostblas.set_threads(threads)
    Cc = np.zeros((n, n), order="F")
    oracle.hostblas.dgemm("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, Cc, n)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.hostblas.dgemm("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, Cc, n)
        times.append(time.perf_counter() - t0)
    cpu_err = lines_error({i: Cc[i, :] for i in rows}, {j: Cc[:, j] for j in cols}, rr, cc)
    flops = 2.0 * n ** 3
    return {"value": round(flops / (gpu["median_ms"] * 1e-3) / 1e12, 3), "unit": "TFLOP/s",
            "config": {"workload": f"C=A*B m=n=k={n} NN alpha=1 beta=0 through cosma::multiply (CosmaMatrix, one "
                                   f"rank, no steps), host matrices in the page-locked memory pool"},
            "gpu": {"ms": gpu["ms"], "median_ms": gpu["median_ms"], "first_call_ms": gpu["ms"][0],
                    "rel_err_sampled_lines_vs_double_double": err},
            "cpu_reference_path": {"tflops": round(flops / min(times) / 1e12, 4),
                                   "min_ms": round(min(times) * 1e3, 3), "threads": threads,
                                   "blas": oracle.hostblas.vendor(),
                                   "rel_err_sampled_lines_vs_double_double": cpu_err},
            "speedup_vs_cpu": round(min(times) * 1e3 / gpu["median_ms"], 2),
            "roofline": {"bound": "host-link", "traffic": 3 * 8.0 * n * n,
                         "note": "96 MB of operands cross the host link per call (A, B up, C down); "
                                 "the 16 GFLOP are 0.2 ms of FP64 MFMA time"}}


def headline(args):
    """The headline DGEMM: its line (without the CPU baseline)."""
    import torch
    import cosma_amd as ca
    say("headline: 16384^3 TN beta=0.5 device-resident")
    dev = torch.device("cuda:0")
    h = ca.Handle(0)
    A = torch.empty((M, K), dtype=torch.float64, device=dev)     # column-major A stored k x m (op(A) = A^T)
    B = torch.empty((N, K), dtype=torch.float64, device=dev)
    C = torch.empty((N, M), dtype=torch.float64, device=dev)
    ca.fill_uniform(A.data_ptr(), K, M, K, 11)
    ca.fill_uniform(B.data_ptr(), K, N, K, 12)
    ca.fill_uniform(C.data_ptr(), M, N, M, 13)
    torch.cuda.synchronize()
    if args.calibrate_copy:
        # 16384 x 16384 doubles = 2 GiB read and 2 GiB written per launch
        h.block_transform("N", M, N, 1.0, B.data_ptr(), K, 0.0, C.data_ptr(), M)
        torch.cuda.synchronize()
        ca.fill_uniform(C.data_ptr(), M, N, M, 13)
    # correctness of one call from the seeded C0 (outside the timed region)
    h.dgemm("T", "N", M, N, K, ALPHA, A.data_ptr(), K, B.data_ptr(), K, BETA, C.data_ptr(), M)
    torch.cuda.synchronize()
    err = None if args.no_check else sampled_error(A, B, C, 13)

    def step():
        h.dgemm("T", "N", M, N, K, ALPHA, A.data_ptr(), K, B.data_ptr(), K, BETA, C.data_ptr(), M)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # the kernel is launched on the null stream, which is torch's current stream
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record()
        step()
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
    flops = 2.0 * M * N * K
    value = flops * args.steps / wall / 1e12
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 uniform[-1,1), generated in HBM)",
        "config": {"workload": "dgemm op(A)=A^T m=n=k=16384 alpha=1 beta=0.5, device-resident",
                   "m": M, "n": N, "k": K, "transa": "T", "transb": "N", "parallelism": "single GPU"},
        "error": {"rel_fro_sampled_rows_vs_double_double": err, "bound_4k_u": 4 * K * 2.0 ** -53},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_MFMA, 4),
                     "traffic": traffic_from_profile()},
    }
    h.close()
    del A, B, C
    torch.cuda.empty_cache()
    return out


def vendor_reference(steps=5, warmup=2):
    """The vendor libraries' GEMMs on this GPU through torch (hipBLASLt /
    rocBLAS), as reference points for the kernels here, not as results: FP64
    at the headline operation (C <- A^T B + 0.5 C, 16384^3) and at 8192^3 NN,
    INT8 x INT8 -> INT32 (torch._int_mm, the INT8 product of the split
    precision) and BF16 at 8192^3 and 16384^3.  Random operands; each entry
    is the best of two timing runs of `steps` calls."""
    import torch
    out = {}

    def run(name, fn, ops, peak, unit):
        t = min(timed(fn, steps, warmup) for _ in range(2))
        rate = ops / t / 1e12
        out[name] = {"ms": round(t * 1e3, 3), "rate": round(rate, 2), "unit": unit, "peak": peak,
                     "frac": round(rate / peak, 4)}
        say(f"vendor {name}: {t * 1e3:.3f} ms, {rate:.1f} {unit}")

    for n in (16384, 8192):
        a = torch.rand(n, n, dtype=torch.float64, device="cuda")
        b = torch.rand(n, n, dtype=torch.float64, device="cuda")
        c = torch.rand(n, n, dtype=torch.float64, device="cuda")
        if n == 16384:
            run("fp64_addmm_tn_16384", lambda: c.addmm_(a.t(), b, beta=BETA, alpha=ALPHA), 2.0 * n ** 3,
                PEAK_FP64_MFMA, "TFLOP/s")
        else:
            run("fp64_matmul_nn_8192", lambda: torch.mm(a, b, out=c), 2.0 * n ** 3, PEAK_FP64_MFMA, "TFLOP/s")
        del a, b, c
        torch.cuda.empty_cache()
    for n in (8192, 16384):
        a = torch.randint(-128, 128, (n, n), dtype=torch.int8, device="cuda")
        b = torch.randint(-128, 128, (n, n), dtype=torch.int8, device="cuda").t()   # k-contiguous columns
        run(f"int8_int_mm_{n}", lambda: torch._int_mm(a, b), 2.0 * n ** 3, PEAK_INT8, "TOP/s")
        del a, b
        x = torch.rand(n, n, dtype=torch.bfloat16, device="cuda")
        y = torch.rand(n, n, dtype=torch.bfloat16, device="cuda")
        run(f"bf16_matmul_{n}", lambda: torch.mm(x, y), 2.0 * n ** 3, PEAK_BF16, "TFLOP/s")
        del x, y
        torch.cuda.empty_cache()
    out["libraries"] = f"torch {torch.__version__} (HIP {torch.version.hip})"
    return out


def leg_line(res, steps, warmup):
    """A leg run alone: its record as a full line of the metric."""
    line = {"metric": METRIC, "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded)"}
    line.update(res)
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    w = args.workload
    if w == "ozaki-target":
        target = split_precision_target(steps=args.steps, warmup=args.warmup)
        print(json.dumps(leg_line({"value": target["moduli"]["tflops_fp64_equiv"], "unit": "TFLOP/s",
                                   "ms_per_step": target["moduli"]["ms"], "config": {"workload": target["workload"]},
                                   "result": target}, args.steps, args.warmup)), flush=True)
        return
    if w == "ozaki":
        res = split_precision(steps=args.steps, warmup=args.warmup)
        best = max(r["moduli"]["tflops_fp64_equiv"] for r in res)
        print(json.dumps(leg_line({"value": best, "unit": "TFLOP/s", "config": {
            "workload": "C=A*B m=n=k=8192, plain FP64 vs split precision, phi in {0,1,4}"}, "results": res},
            args.steps, args.warmup)), flush=True)
        return
    if w == "ozaki-8192":
        res = split_precision(phis=(args.phi,), steps=args.steps, warmup=args.warmup, sweep=False, methods=("moduli",))
        r = res[0]
        print(json.dumps(leg_line({"value": r["moduli"]["tflops_fp64_equiv"], "unit": "TFLOP/s",
                                   "ms_per_step": r["moduli"]["ms"], "config": {
                                       "workload": f"C=A*B m=n=k=8192, moduli method at 1e-14, phi={args.phi}"},
                                   "result": r}, args.steps, args.warmup)), flush=True)
        return
    if w in ("host", "tall-skinny"):
        print(json.dumps(leg_line(streamed(w, args.steps, args.warmup, not args.no_check), args.steps,
                                  args.warmup)), flush=True)
        return
    if w == "reorder":
        print(json.dumps(leg_line(reorder(args.steps, args.warmup), args.steps, args.warmup)), flush=True)
        return
    if w == "batched":
        print(json.dumps(leg_line(batched(args.steps, args.warmup), args.steps, args.warmup)), flush=True)
        return
    if w == "miniapp":
        print(json.dumps(leg_line(miniapp(), args.steps, args.warmup)), flush=True)
        return
    if w == "vendor":
        res = vendor_reference(args.steps, args.warmup)
        print(json.dumps(leg_line({"value": res["fp64_addmm_tn_16384"]["rate"], "unit": "TFLOP/s", "config": {
            "workload": "torch GEMMs on this GPU (reference points)"}, "vendor_reference": res},
            args.steps, args.warmup)), flush=True)
        return

    out = headline(args)
    if not args.no_split:
        out["split_precision"] = split_precision_target()
    if not args.quick and w is None:
        t0 = time.perf_counter()
        out["configs"] = {
            "1_miniapp_2000": miniapp(),
            "2_host_resident_16384": streamed("host", 3, 1, not args.no_check),
            "2_reorder_16384": reorder(5, 2),
            "3_tall_skinny_2048x2048x1048576": streamed("tall-skinny", 2, 1, not args.no_check),
            "4_split_precision_8192": split_precision(steps=3, warmup=1, sweep=False),
            "5_batched_4096": batched(10, 2),
        }
        out["vendor_reference"] = vendor_reference()
        out["configs_wall_s"] = round(time.perf_counter() - t0, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
