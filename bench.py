"""Benchmark of the AP26 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain|prp]
                    [--no-cpu] [--variant N]

Prints ONE JSON line.  A step is one searched K of the plain single-GPU
workload: the full 10-shift search of that K (tables -> sieve -> SPRP walk ->
AP check) plus the host-side ReportSolution/hash of its reports.  The K are the
consecutive K from 48620635 (numpy default_rng(26).integers(4e7, 5e7+1)) at
SHIFT 640 that pass will_search; skipped K are not steps.  Each step copies
the K's 10840 n43 starts (86 KB) host->device, as the reference's plain GPU
path does; everything else is generated on the device.

Extra objects on the line:
  roofline      the sieve kernel (the dominant one) priced on the INT32 VALU
                roofline from algorithmic counts, timed live with HIP events;
  prp           the isolated SPRP micro-benchmark (2^26 odd values in
                [2^58, 2^63)) and its INT32 roofline;
  cpu_baseline  the oracle (a from-scratch C restatement of the reference CPU
                search) on a bounded sample of the same workload, all host
                cores the box grants (up to 16).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ap26_amd import gpu  # noqa: E402
from ap26_amd import host as H  # noqa: E402

METRIC = "K/sec (10-shift) + PRP tests/sec, batched vs plain GPU; p95 work-unit latency"
K0_C3 = 48620635          # default_rng(26).integers(40_000_000, 50_000_001)
SHIFT_C3 = 640
N59_PER_K = 10840 * 12673 * 35  # positions per shift per K
# Lookups per 64-bit mask evaluation under the reference GPU kernel's
# early-exit grouping (src/opencl/kernels/sieve.cl:32-114), measured with the
# oracle's ap26o_lookup_count over 40 n43 starts spread across K=48620635,
# SHIFT 640 (all 10 shifts): 19.466.
LOOKUPS_PER_MASK_C3 = 19.466
# INT32 op-equivalents per lookup: quotient estimate + multiply-subtract
# (residue), AND of the 64-bit mask (2).  The table load itself is not counted.
OPS_PER_LOOKUP = 4
# 64-bit Montgomery squaring: 18 MUL32 + 16 add-class INT32 ops.
OPS_PER_SQUARING = 34
PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.3 T int32 lane-ops/s (nominal)


def will_search(K):
    return all(K % p for p in (29, 31, 37, 41, 43, 47, 53, 59))


def c3_ks(n, start=0):
    out = []
    K = K0_C3
    while len(out) < start + n:
        if will_search(K):
            out.append(K)
        K += 1
    return out[start:]


def run_plain(ctx, steps, warmup):
    ks = c3_ks(warmup + steps)
    tmp = tempfile.mkdtemp(prefix="ap26bench")
    h = H.Host(ks[0], ks[-1], SHIFT_C3, tmp, min_report=20)
    for K in ks[:warmup]:
        sols, _ = ctx.search_k(K, SHIFT_C3)
    ctx.sync()
    sieve_ms = 0.0
    launches = 0
    positions = 0
    surv = 0
    tests = 0
    walk_ms = 0.0
    per_k = []
    t0 = time.perf_counter()
    for K in ks[warmup:]:
        tk = time.perf_counter()
        sols, st = ctx.search_k(K, SHIFT_C3)
        for s in sorted(sols.tolist(), key=lambda s: (s[2], s[0])):
            if h.report(s[0], s[1], s[2]) != 0:
                raise RuntimeError("report failed")
        per_k.append(time.perf_counter() - tk)
        sieve_ms += st.ms_sieve
        walk_ms += st.ms_walk
        launches += st.sieve_launches
        positions += st.positions
        surv += st.survivors
        tests += st.prp_tests
    ctx.sync()
    dt = time.perf_counter() - t0
    h.close()
    return {"seconds": dt, "ks": ks[warmup:], "sieve_ms": sieve_ms, "walk_ms": walk_ms,
            "launches": launches, "positions": positions, "survivors": surv, "prp_tests": tests,
            "per_k_s": per_k, "hash": h.hash}


def run_prp(ctx, count=1 << 26, seed=26):
    rng = np.random.default_rng(seed)
    n = rng.integers(2 ** 58, 2 ** 63, size=count, dtype=np.uint64) | np.uint64(1)
    # W(N) = bitlen(N) - 2 squarings, counted exactly on the inputs
    bl = np.floor(np.log2(n.astype(np.float64))).astype(np.int64) + 1
    squarings = int((bl - 2).sum())
    d_n = ctx.dev_alloc(n.nbytes)
    d_o = ctx.dev_alloc(count)
    ctx.h2d(d_n, n)
    ctx.prp_batch_dev(d_n, d_o, count)  # warm-up
    times = [ctx.prp_batch_dev(d_n, d_o, count) for _ in range(5)]
    out = np.zeros(count, dtype=np.uint8)
    ctx.d2h(out, d_o)
    ctx.dev_free(d_n)
    ctx.dev_free(d_o)
    ms = float(np.median(times))
    ops = squarings * OPS_PER_SQUARING
    return {"tests": count, "ms": ms, "tests_per_s": count / (ms / 1e3),
            "int32_tops": ops / (ms / 1e3) / 1e12, "prp_count": int(out.sum()),
            "mean_squarings": squarings / count}


def cpu_baseline(threads, j1=2048):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    K = K0_C3
    t = time.perf_counter()
    O.search_span(K, SHIFT_C3, 0, 10, 0, j1, threads=threads)
    dt = time.perf_counter() - t
    frac = j1 / 10840
    return {"value": frac / dt, "unit": "K/s", "cores": threads, "kind": "port",
            "sample": f"oracle search of K={K} SHIFT {SHIFT_C3}, all 10 shifts, n43 starts "
                      f"0..{j1 - 1} ({frac * 100:.1f}% of one K) in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-prp", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the search shards by K with no inter-GPU traffic; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    ctx = gpu.Context(0, sieve_variant=args.variant)
    r = run_plain(ctx, args.steps, args.warmup)
    kps = len(r["ks"]) / r["seconds"]
    skipped = (r["ks"][-1] - r["ks"][0] + 1) - len(r["ks"])
    avg_launch_s = r["sieve_ms"] / r["launches"] / 1e3
    pos_per_launch = r["positions"] / r["launches"]
    ops_per_launch = pos_per_launch * LOOKUPS_PER_MASK_C3 * OPS_PER_LOOKUP
    achieved = ops_per_launch / avg_launch_s / 1e12
    line = {
        "metric": METRIC,
        "value": kps,
        "unit": "K/s",
        "n_gpus": 1,
        "steps": len(r["ks"]),
        "warmup": args.warmup,
        "ms_per_step": r["seconds"] * 1e3 / len(r["ks"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "plain single-GPU path: full 10-shift search of consecutive K from "
                               f"{K0_C3}, SHIFT {SHIFT_C3}, one kernel sequence per K",
                   "K_first": r["ks"][0], "K_last": r["ks"][-1], "k_searched": len(r["ks"]),
                   "k_skipped": skipped, "shift": SHIFT_C3, "sieve_variant": args.variant,
                   "p95_k_latency_ms": float(np.percentile(r["per_k_s"], 95) * 1e3)},
        "roofline": {"bound": "valu_int32", "achieved": achieved, "peak": PEAK_INT32_TOPS,
                     "unit": "TOP/s", "frac": achieved / PEAK_INT32_TOPS, "traffic": None,
                     "kernel": "sieve_queue_kernel" if args.variant == 0 else "sieve_kernel",
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "ops_per_launch": ops_per_launch,
                     "basis": f"{LOOKUPS_PER_MASK_C3} lookups/mask x {OPS_PER_LOOKUP} int32 ops "
                              f"x {pos_per_launch:.4g} masks/launch"},
        "detail": {"sieve_ms_per_k": r["sieve_ms"] / len(r["ks"]),
                   "walk_ms_per_k": r["walk_ms"] / len(r["ks"]),
                   "survivors_per_k": r["survivors"] / len(r["ks"]),
                   "prp_tests_per_k": r["prp_tests"] / len(r["ks"])},
    }
    if not args.no_prp:
        p = run_prp(ctx)
        line["prp"] = {"tests_per_s": p["tests_per_s"], "ms": p["ms"], "tests": p["tests"],
                       "mean_squarings": p["mean_squarings"],
                       "roofline": {"bound": "valu_int32", "achieved": p["int32_tops"],
                                    "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                                    "frac": p["int32_tops"] / PEAK_INT32_TOPS}}
    ctx.close()
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        line["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
