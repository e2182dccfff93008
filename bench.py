"""Benchmark of the single-rank local multiply on one MI355X.

Workload (the plain-FP64 configuration the headline metric is quoted on):
C <- alpha * A^T * B + beta * C with m = n = k = 16384, FP64, A, B, C uniform
[-1, 1) (seeded SplitMix64, generated in HBM), alpha = 1, beta = 0.5,
operands device-resident.  One step = one DGEMM call through the C ABI.

Prints one JSON line with the throughput (TFLOP/s = 2mnk / t), the measured
roofline fraction of the DGEMM kernel (FP64 MFMA peak), the error of the
result against the double-double oracle on sampled rows, and the reference's
CPU path (one host-BLAS cblas_dgemm, src/cosma/blas.cpp:35) timed on a
bounded sample on this machine's cores.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X datasheet dense FP64 matrix
M = N = K = 16384
ALPHA, BETA = 1.0, 0.5
METRIC = "FP64-accurate TFLOP/s, 1 GPU: native DGEMM vs split-precision at stated error"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="device",
                   choices=["device", "host", "tall-skinny", "ozaki", "ozaki-target", "reorder", "batched"],
                   help="device: 16384^3 TN on device-resident operands (default, the headline); "
                        "host: same GEMM on pinned host operands streamed in 4096^2 tiles; "
                        "tall-skinny: m=n=2048, k=1048576 A^T B streamed from pinned host memory "
                        "in k-panels of 65536; ozaki: 8192^3 split-precision vs plain FP64 for "
                        "phi in {0, 1, 4} plus the 16384^3 A^T B comparison; ozaki-target: only "
                        "that 16384^3 comparison; reorder: the layout transforms of a 16384^2 pdgemm "
                        "(A transposed, B copied, C <- alpha X + beta C) on the device; "
                        "batched: 4096 NN problems with m, n, k log-uniform in [16, 512] in one "
                        "grouped launch replayed as a hipGraph")
    p.add_argument("--no-split", action="store_true", help="skip the split-precision summary leg")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    p.add_argument("--cpu-n", type=int, default=8192, help="CPU sample size (m = n = k)")
    p.add_argument("--no-check", action="store_true", help="skip the sampled-row error check")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a 2 GiB device copy (block_transform 'N', 2 GiB read + 2 GiB "
                        "written) to calibrate HBM byte counters")
    return p.parse_args()


def traffic_from_profile():
    """HBM bytes per launch of the DGEMM kernel from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "dgemm_tn_16384_pmc.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def cpu_baseline(n):
    """The reference CPU path (one host-BLAS dgemm, op(A) = A^T, beta = 0.5) on n^3."""
    import numpy as np
    import oracle
    threads = min(16, len(os.sched_getaffinity(0)))
    oracle.hostblas.set_threads(threads)
    A, B, C = oracle.uniform(n, n, 1), oracle.uniform(n, n, 2), oracle.uniform(n, n, 3)
    oracle.hostblas.dgemm("T", "N", n, n, n, ALPHA, A, n, B, n, BETA, C, n)   # warm-up
    times = []
    budget = time.perf_counter() + 25.0
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.hostblas.dgemm("T", "N", n, n, n, ALPHA, A, n, B, n, BETA, C, n)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > budget:
            break
    best = min(times)
    return {"value": round(2.0 * n ** 3 / best / 1e12, 4), "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"cblas_dgemm('T','N') m=n=k={n}, alpha=1, beta=0.5, uniform[-1,1) inputs, "
                      f"min of {len(times)} reps after 1 warm-up, BLAS {oracle.hostblas.vendor()}"}


def sampled_error(A, B, C, C0_seed, rows=(0, 8191)):
    """Relative error of sampled rows of C against the double-double oracle."""
    import numpy as np
    import oracle
    Bh = np.asfortranarray(B.cpu().numpy().T)
    num = den = 0.0
    for i in rows:
        a_row = np.asfortranarray(A[i].cpu().numpy().reshape(K, 1))
        ref = np.zeros((1, N), order="F")
        oracle.dgemm("T", "N", 1, N, K, ALPHA, a_row, K, Bh, K, 0.0, ref, 1, dd=True)
        c0 = oracle.uniform_at(C0_seed, i + np.arange(N, dtype=np.uint64) * M)
        expect = ref[0] + BETA * c0
        got = C[:, i].cpu().numpy()
        num += float(((got - expect) ** 2).sum())
        den += float((expect ** 2).sum())
    return (num / den) ** 0.5


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


def streamed_error(h, ta, tb, m, n, k, alpha, beta, host, C0_seed):
    """One more call from the seeded C0, then the relative error of rows 0 and
    m-1 and column n/2 of C against the double-double oracle (all k terms)."""
    import torch
    import cosma_amd as ca
    from oracle.sampled import dd_lines, lines_error
    C = host["C"]                                   # torch (n, m): row j = column j of C
    cols_dev = max(1, (1 << 28) // m)
    tmp = torch.empty((min(cols_dev, n), m), dtype=torch.float64, device="cuda")
    for c0 in range(0, n, cols_dev):
        cc = min(cols_dev, n - c0)
        ca.fill_uniform(tmp.data_ptr(), m, cc, m, C0_seed, offset=c0 * m)
        C[c0:c0 + cc].copy_(tmp[:cc])
    torch.cuda.synchronize()
    del tmp
    rows, cols = (0, m - 1), (n // 2,)
    c0r = {i: C[:, i].numpy().copy() for i in rows}
    c0c = {j: C[j, :].numpy().copy() for j in cols}
    h.dgemm_host(ta, tb, m, n, k, alpha, host["A"].data_ptr(), k, host["B"].data_ptr(), k, beta, C.data_ptr(), m)
    r, c = dd_lines(ta, tb, m, n, k, host["A"].data_ptr(), k, host["B"].data_ptr(), k, rows, cols)
    return lines_error({i: C[:, i].numpy() for i in rows}, {j: C[j, :].numpy() for j in cols},
                       {i: alpha * r[i] + beta * c0r[i] for i in rows},
                       {j: alpha * c[j] + beta * c0c[j] for j in cols})


def streamed(args):
    """Host-resident workloads through the host-streaming DGEMM."""
    import torch
    import cosma_amd as ca
    if args.workload == "host":
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
    h = ca.Handle(0, streams=2, tile_m=tiles[0], tile_n=tiles[1], tile_k=tiles[2])
    # pinned host operands (column-major; A stored k x m), filled on the GPU and copied down
    shapes = {"A": (m, k), "B": (n, k), "C": (n, m)}            # torch (cols, rows)
    host = {x: torch.empty(sh, dtype=torch.float64, pin_memory=True) for x, sh in shapes.items()}
    for seed, (x, sh) in enumerate(shapes.items(), start=11):
        cols = sh[0]
        step_cols = max(1, (1 << 28) // sh[1])
        tmp = torch.empty((min(step_cols, cols), sh[1]), dtype=torch.float64, device="cuda")
        for c0 in range(0, cols, step_cols):
            cc = min(step_cols, cols - c0)
            ca.fill_uniform(tmp.data_ptr(), sh[1], cc, sh[1], seed, offset=c0 * sh[1])
            host[x][c0:c0 + cc].copy_(tmp[:cc])
        del tmp
    torch.cuda.synchronize()
    C0_seed = 13
    link = link_bandwidth(torch)

    def call():
        h.dgemm_host(ta, tb, m, n, k, alpha, host["A"].data_ptr(), k, host["B"].data_ptr(), k, beta,
                     host["C"].data_ptr(), m)

    times = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        call()
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            times.append(dt)
    t = sum(times) / len(times)
    err = None if args.no_check else streamed_error(h, ta, tb, m, n, k, alpha, beta, host, C0_seed)
    flops = 2.0 * m * n * k
    h2d = 8.0 * (m * k + k * n + (m * n if beta != 0 else 0))
    d2h = 8.0 * m * n
    t_mfma = flops / (PEAK_FP64_MFMA * 1e12)
    t_link = h2d / (link * 1e9)
    out = {
        "metric": METRIC, "value": round(flops / t / 1e12, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 uniform[-1,1))",
        "config": {"workload": desc, "m": m, "n": n, "k": k, "tiles": tiles, "schedule":
                   {0: "device", 1: "resident panels", 2: "bounded ring"}[h.last_mode()]},
        "error": {"rel_fro_sampled_lines_vs_double_double": err, "bound_4k_u": 4 * k * 2.0 ** -53,
                  "lines": "rows 0, m-1 and column n/2 of C, each with all k terms"},
        "roofline": {"bound": "host-link" if t_link > t_mfma else "mfma",
                     "achieved": round(h2d / t / 1e9, 2), "peak": round(link, 2), "unit": "GB/s",
                     "frac": round(max(t_link, t_mfma) / t, 4), "traffic": h2d + d2h,
                     "note": "achieved = host->device bytes / time; peak = measured pinned H2D copy rate; "
                             "frac = max(link time, MFMA time) / time"},
    }
    print(json.dumps(out), flush=True)


def rows_error(Ad, Bd, Cd, n, rows, ref_rows):
    """Relative Frobenius error of sampled rows of C (device, column-major n x n)."""
    num = den = 0.0
    for i, ref in zip(rows, ref_rows):
        got = Cd[:, i].cpu().numpy()
        num += float(((got - ref) ** 2).sum())
        den += float((ref ** 2).sum())
    return (num / den) ** 0.5


def dd_rows(Ad, Bd, n, rows):
    """Rows of A*B (NN, n x n column-major device operands) from the double-double oracle."""
    import numpy as np
    import oracle
    Bh = np.asfortranarray(Bd.cpu().numpy().T)
    Ah = Ad.cpu().numpy()                 # tensor (k, m): row l = column l of A
    out = []
    for i in rows:
        a_row = np.asfortranarray(Ah[:, i].reshape(1, n))
        ref = np.zeros((1, n), order="F")
        oracle.dgemm("N", "N", 1, n, n, 1.0, a_row, 1, Bh, n, 0.0, ref, 1, dd=True)
        out.append(ref[0])
    return out


def split_precision(phis=(0.0, 1.0, 4.0), steps=5, warmup=1, n=8192, rows=(0, 4097)):
    """8192^3 C = A*B (NN): plain FP64 MFMA DGEMM vs the split-precision DGEMM
    (moduli at the default 1e-14 tolerance, with INT8 and with BF16 residues;
    slices with 7 slices), inputs
    a = (u - 0.5) exp(phi z).  Errors are relative Frobenius errors of sampled
    rows against the double-double oracle."""
    import torch
    import cosma_amd as ca
    h = ca.Handle(0)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    B = torch.empty((n, n), dtype=torch.float64, device="cuda")
    C = torch.empty((n, n), dtype=torch.float64, device="cuda")
    flops = 2.0 * n ** 3
    results = []
    for phi in phis:
        ca.fill_lognormal(A.data_ptr(), n, n, n, 21, phi)
        ca.fill_lognormal(B.data_ptr(), n, n, n, 22, phi)
        torch.cuda.synchronize()
        ref = dd_rows(A, B, n, rows)

        def timed(fn):
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / steps

        plain = lambda: h.dgemm("N", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n)
        t_plain = timed(plain)
        e_plain = rows_error(A, B, C, n, rows, ref)
        entry = {"phi": phi, "fp64": {"tflops": round(flops / t_plain / 1e12, 3), "ms": round(t_plain * 1e3, 3),
                                      "rel_err": e_plain}}
        for name, method, count in (("moduli", ca.OZAKI_MODULI, 0), ("moduli_bf16", ca.OZAKI_MODULI_BF16, 0),
                                    ("slices", ca.OZAKI_SLICES, 7)):
            info = {}

            def oz():
                info.update(h.dgemm_ozaki("N", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0,
                                          C.data_ptr(), n, method=method, tolerance=1e-14, count=count))
            t = timed(oz)
            entry[name] = {"tflops_fp64_equiv": round(flops / t / 1e12, 3), "ms": round(t * 1e3, 3),
                           "speedup_vs_fp64": round(t_plain / t, 3), "rel_err": rows_error(A, B, C, n, rows, ref),
                           "count": info["count"], "bits": info["bits"], "int8_products": info["int8_products"],
                           "error_estimate": info["error_bound"]}
        # error against the double-double oracle for each count: moduli 8..20
        # and slices 4..8 (one warm-up and one timed call each)
        sweep = []
        for name, method, counts in (("moduli", ca.OZAKI_MODULI, range(8, 21, 2)),
                                     ("slices", ca.OZAKI_SLICES, range(4, 9))):
            for cnt in counts:
                info = {}

                def oz():
                    info.update(h.dgemm_ozaki("N", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0,
                                              C.data_ptr(), n, method=method, count=cnt))
                oz()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                oz()
                torch.cuda.synchronize()
                t = time.perf_counter() - t0
                sweep.append({"method": name, "count": cnt, "bits": info["bits"], "ms": round(t * 1e3, 3),
                              "rel_err": rows_error(A, B, C, n, rows, ref), "error_estimate": info["error_bound"]})
        entry["count_sweep"] = sweep
        results.append(entry)
    del A, B, C
    torch.cuda.empty_cache()
    return results


def dd_rows_tn(Ad, Bd, n, rows):
    """Rows of A^T*B (A stored k x m, B k x n, column-major device operands)
    from the double-double oracle."""
    import numpy as np
    import oracle
    Bh = np.asfortranarray(Bd.cpu().numpy().T)
    Ah = Ad.cpu().numpy()                 # tensor (m, k): row i = column i of stored A = row i of A^T
    out = []
    for i in rows:
        a_row = np.asfortranarray(Ah[i, :].reshape(1, n))
        ref = np.zeros((1, n), order="F")
        oracle.dgemm("N", "N", 1, n, n, 1.0, a_row, 1, Bh, n, 0.0, ref, 1, dd=True)
        out.append(ref[0])
    return out


def split_precision_target(steps=3, warmup=1, n=16384, rows=(0, 8193)):
    """Split precision at the plain FP64 shape: m = n = k = 16384, op(A) = A^T,
    alpha = 1, uniform [-1, 1) inputs, beta = 0 (so that C needs no restoring
    between timed calls): plain FP64 MFMA DGEMM vs the moduli method at the
    default 1e-14 tolerance, errors of sampled rows against the double-double
    oracle."""
    import torch
    import cosma_amd as ca
    h = ca.Handle(0)
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    B = torch.empty((n, n), dtype=torch.float64, device="cuda")
    C = torch.empty((n, n), dtype=torch.float64, device="cuda")
    ca.fill_uniform(A.data_ptr(), n, n, n, 41)
    ca.fill_uniform(B.data_ptr(), n, n, n, 42)
    torch.cuda.synchronize()
    ref = dd_rows_tn(A, B, n, rows)
    flops = 2.0 * n ** 3

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_plain = timed(lambda: h.dgemm("T", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n))
    e_plain = rows_error(A, B, C, n, rows, ref)
    info = {}

    def oz():
        info.update(h.dgemm_ozaki("T", "N", n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, C.data_ptr(), n,
                                  method=ca.OZAKI_MODULI, tolerance=1e-14))
    t_oz = timed(oz)
    out = {"workload": "C=A^T*B m=n=k=16384 alpha=1 beta=0, uniform[-1,1) inputs",
           "fp64": {"tflops": round(flops / t_plain / 1e12, 3), "ms": round(t_plain * 1e3, 3), "rel_err": e_plain},
           "moduli": {"tflops_fp64_equiv": round(flops / t_oz / 1e12, 3), "ms": round(t_oz * 1e3, 3),
                      "speedup_vs_fp64": round(t_plain / t_oz, 3), "rel_err": rows_error(A, B, C, n, rows, ref),
                      "count": info["count"], "bits": info["bits"], "int8_products": info["int8_products"],
                      "error_estimate": info["error_bound"]}}
    del A, B, C
    torch.cuda.empty_cache()
    return out


PEAK_HBM = 8000.0  # GB/s, MI355X datasheet


def reorder(args):
    """Device block transforms of the ScaLAPACK -> COSMA -> ScaLAPACK round trip
    of a 16384^2 pdgemm on one rank: A transposed (op = 'T'), B copied
    (op = 'N'), the result written back as C <- alpha * X + beta * C."""
    import torch
    import cosma_amd as ca
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
    for name, (fn, nbytes) in legs.items():
        for _ in range(args.warmup):
            fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        s.record()
        for _ in range(args.steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / args.steps
        res[name] = {"ms": round(ms, 4), "GB/s": round(nbytes / ms / 1e6, 1),
                     "frac_of_hbm": round(nbytes / ms / 1e6 / PEAK_HBM, 4)}
        total_b += nbytes
        total_t += ms
    out = {"metric": "device layout-transform bandwidth (algorithmic bytes / time)",
           "value": round(total_b / total_t / 1e6, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(total_t, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded uniform)",
           "config": {"workload": "16384^2 ScaLAPACK <-> COSMA transforms of one pdgemm on one rank"},
           "legs": res,
           "roofline": {"bound": "hbm", "achieved": round(total_b / total_t / 1e6, 1), "peak": PEAK_HBM,
                        "unit": "GB/s", "frac": round(total_b / total_t / 1e6 / PEAK_HBM, 4), "traffic": None}}
    print(json.dumps(out), flush=True)


def batch_dims(count=4096, seed=5):
    """m, n, k of the batched workload: each round(exp(U[ln 16, ln 512]))
    clamped to [16, 512], numpy default_rng(seed); shape (count, 3)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.clip(np.rint(np.exp(rng.uniform(np.log(16), np.log(512), (count, 3)))), 16, 512).astype(np.int64)


def batched(args, count=4096, seed=5):
    """4096 independent C = A B problems (NN, alpha = 1, beta = 0), m, n, k each
    round(exp(U[ln 16, ln 512])) clamped to [16, 512] (numpy default_rng(5)),
    operands packed in three device buffers; one grouped persistent launch per
    step, replayed as a hipGraph after the plan is tuned; then the same batch
    in split-precision form (grouped moduli plan at the 1e-14 tolerance)."""
    import numpy as np
    import torch
    import cosma_amd as ca
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
    for _ in range(args.warmup):
        plan.graph_run()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(args.steps):
        plan.graph_run()
    e.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = s.elapsed_time(e) / args.steps
    flops = 2.0 * (m * n * k).astype(np.float64)
    nbytes = 8.0 * (m * k + k * n + m * n).astype(np.float64)
    t_roof = float(np.maximum(flops / (PEAK_FP64_MFMA * 1e12), nbytes / (PEAK_HBM * 1e9)).sum())
    # correctness of a sample of problems against the double-double oracle
    import oracle
    refs = {}

    def sampled_error():
        errs = []
        for i in (0, 1, count // 2, count - 1):
            Ch = C[int(oc[i]):int(oc[i]) + int(m[i] * n[i])].cpu().numpy().reshape((int(n[i]), int(m[i]))).T
            if i not in refs:
                Ah = A[int(oa[i]):int(oa[i]) + int(m[i] * k[i])].cpu().numpy().reshape((int(k[i]), int(m[i]))).T
                Bh = B[int(ob[i]):int(ob[i]) + int(k[i] * n[i])].cpu().numpy().reshape((int(n[i]), int(k[i]))).T
                ref = np.zeros((int(m[i]), int(n[i])), order="F")
                oracle.dgemm("N", "N", int(m[i]), int(n[i]), int(k[i]), 1.0, np.asfortranarray(Ah), int(m[i]),
                             np.asfortranarray(Bh), int(k[i]), 0.0, ref, int(m[i]), dd=True)
                refs[i] = ref
            errs.append(oracle.rel_fro(Ch, refs[i]))
        return max(errs)

    errs = [sampled_error()]
    info = plan.info()
    value = float(flops.sum()) / wall / 1e12
    out = {"metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded uniform operands, seeded log-uniform sizes)",
           "config": {"workload": f"{count} NN problems, m,n,k log-uniform in [16,512], one grouped launch per "
                                  f"step replayed as a hipGraph", "sum_flops": float(flops.sum()),
                      "sum_bytes": float(nbytes.sum()), "plan": info},
           "error": {"max_rel_fro_sampled_problems_vs_double_double": max(errs)},
           "roofline": {"bound": "per-problem max(MFMA, HBM)", "achieved": round(value, 3),
                        "peak": round(float(flops.sum()) / t_roof / 1e12, 3), "unit": "TFLOP/s",
                        "frac": round(t_roof / (ms * 1e-3), 4), "traffic": None,
                        "note": "peak = sum flops / sum_i max(flops_i / 78.6 TF, bytes_i / 8 TB/s); "
                                "frac uses the launch-inclusive event time of a graph replay"}}
    plan.close()
    # the split-precision form of the same batch: one plan at the default 1e-14
    # tolerance (moduli chosen from these operands), tuned, replayed as a graph
    oz = ca.GroupedOzaki(h, probs)
    oz.tune()
    for _ in range(args.warmup):
        oz.graph_run()
    torch.cuda.synchronize()
    s.record()
    for _ in range(args.steps):
        oz.graph_run()
    e.record()
    torch.cuda.synchronize()
    oz_ms = s.elapsed_time(e) / args.steps
    out["ozaki_form"] = {"ms_per_step": round(oz_ms, 4),
                         "tflops_fp64_equiv": round(float(flops.sum()) / (oz_ms * 1e-3) / 1e12, 3),
                         "vs_plain_fp64": round(ms / oz_ms, 3), "plan": oz.info(),
                         "max_rel_fro_sampled_problems_vs_double_double": sampled_error()}
    oz.close()
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "ozaki-target":
        target = split_precision_target(steps=args.steps, warmup=args.warmup)
        out = {"metric": METRIC, "value": target["moduli"]["tflops_fp64_equiv"], "unit": "TFLOP/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": target["moduli"]["ms"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 (int8 MFMA inside)",
               "data": "synthetic (seeded SplitMix64 uniform[-1,1), generated in HBM)",
               "config": {"workload": target["workload"]}, "result": target}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "ozaki":
        res = split_precision(steps=args.steps, warmup=args.warmup)
        target = split_precision_target(steps=min(args.steps, 3), warmup=1)
        best = max(r["moduli"]["tflops_fp64_equiv"] for r in res)
        out = {"metric": METRIC, "value": best, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64 (int8 MFMA inside)", "data": "synthetic a=(u-0.5)exp(phi z), seeded",
               "config": {"workload": "C=A*B m=n=k=8192, plain FP64 vs split precision, phi in {0,1,4}"},
               "results": res, "target_shape": target,
               "error_note": "rel_err = relative Frobenius error of 2 sampled rows vs the double-double oracle"}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "reorder":
        return reorder(args)
    if args.workload == "batched":
        return batched(args)
    if args.workload != "device":
        return streamed(args)
    import torch
    import cosma_amd as ca

    dev = torch.device("cuda:0")
    h = ca.Handle(0)
    # column-major 16384 x 16384 operands; A is stored k x m (op(A) = A^T)
    A = torch.empty((M, K), dtype=torch.float64, device=dev)
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
    del A, B, C
    torch.cuda.empty_cache()
    if not args.no_split:
        out["split_precision"] = split_precision_target()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
