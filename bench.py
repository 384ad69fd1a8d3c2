"""Benchmark of the AP26 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-steps N]
                    [--no-cpu] [--no-serve] [--no-plain]

Prints ONE JSON line.  The headline `value` is K/sec (10-shift) of the batched
persistent path; a step is one searched K.  The timed K are a bounded sample
of the batched workload: 256 units from numpy default_rng(26) with length
floor(exp(U(0, ln 4096))), start uniform in [1e6, 1e8], SHIFT from
{0, 640, 1280}; the sample takes the first min(length, 4) K of units in order
until `steps` searched K are collected, so it mixes SHIFTs, unit lengths and
K magnitudes.  The whole unit is searched -- tables, all 10 shifts of the
sieve, the SPRP walk, AP check, per-range hash totals on the device -- and the
host applies ReportSolution (9-base check, hash) to every report.

The batched engine's sieve variant and work-item size are chosen first by the
library's autotuner (ap26_autotune, 10 candidates timed on K=48620637).

Beside it on the line:
  plain         the plain per-K path (one kernel sequence per K, host sync per
                K) on consecutive K from 48620635 at SHIFT 640;
  roofline      the batched path's persistent sieve kernel (the dominant one)
                priced on the INT32 VALU roofline from algorithmic counts,
                timed with HIP events recorded inside the replayed graph;
  prp           the isolated SPRP micro-benchmark (2^26 odd values in
                [2^58, 2^63)) and its INT32 roofline;
  serving       continuous batching with Poisson arrivals at half the
                saturated unit rate, unit length floor(exp(U(0, ln 16))) (the
                workload's 1..8192 is scaled down so that units finish within
                the bench), p50/p95/p99 unit latency;
  cpu_baseline  the oracle (a from-scratch C restatement of the reference CPU
                search) on a bounded sample of the plain workload, on the host
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
MASKS_PER_SHIFT = 10840 * 12673 * 35  # n59 positions per shift per K
# Lookups per 64-bit mask evaluation under the reference GPU kernel's
# early-exit grouping (src/opencl/kernels/sieve.cl:32-114), measured with the
# oracle's ap26o_lookup_count over 40 n43 starts spread across K=48620635 at
# SHIFT 640 (all 10 shifts): 19.466; over K=366384 at SHIFT 0: 18.224; over
# K=875859 at SHIFT 640: 18.423.  The batched sample mixes K from 1e6 to 1e8
# and uses the mean of the three.
LOOKUPS_PER_MASK_C3 = 19.466
LOOKUPS_PER_MASK_MIX = (19.466 + 18.224 + 18.423) / 3
# INT32 op-equivalents per lookup: quotient estimate + multiply-subtract
# (residue), AND of the 64-bit mask (2).  The table load itself is not counted.
OPS_PER_LOOKUP = 4
# 64-bit Montgomery squaring: 18 MUL32 + 16 add-class INT32 ops.
OPS_PER_SQUARING = 34
PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.3 T int32 lane-ops/s (nominal)
PEAK_HBM_GBS = 8000.0
# HBM bytes per searched K moved by the persistent sieve kernel, from two
# rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over a 16-K round of this
# workload (profiles/sieve_persistent2_traffic.md): FETCH_SIZE 291,960 KiB and
# WRITE_SIZE 15,961,158 KiB per launch.  Both are scaled by what the same
# counters read for prp_kernel over 2^26 values, which reads 512 MiB and
# writes 64 MiB exactly: FETCH_SIZE 262,174 KiB (x2.000), WRITE_SIZE
# 65,536 KiB (x1.000).  (2 x 291,960 + 15,961,158) KiB / 16 K:
TRAFFIC_BYTES_PER_K = (2.0 * 291960.4 + 15961158.4) * 1024 / 16
# Algorithmic bytes of the sieve: each survivor written once (8-byte value +
# 4-byte job index); the per-K tables (10 shifts x 23,783 words x 8 B) and
# n43 starts (10,840 x 8 B) read once.
SURVIVOR_BYTES = 12
TABLE_BYTES_PER_K = 10 * 23783 * 8 + 10840 * 8


def will_search(K):
    return all(K % p for p in (29, 31, 37, 41, 43, 47, 53, 59))


def c3_ks(n):
    out, K = [], K0_C3
    while len(out) < n:
        if will_search(K):
            out.append(K)
        K += 1
    return out


def c4_units(seed=26, n=256):
    rng = np.random.default_rng(seed)
    lens = np.clip(np.floor(np.exp(rng.uniform(0, np.log(4096), n))).astype(int), 1, 4096)
    starts = rng.integers(1_000_000, 100_000_001, n)
    shifts = rng.choice([0, 640, 1280], n)
    return [(int(s), int(s + L - 1), int(sh)) for s, L, sh in zip(starts, lens, shifts)]


def c4_sample(steps, per_unit=4):
    """Units (kmin, kmax, shift, 0) covering the first `steps` searched K."""
    out, got = [], 0
    for kmin, kmax, sh in c4_units():
        ks = [K for K in range(kmin, kmax + 1) if will_search(K)][:per_unit]
        if not ks:
            continue
        ks = ks[:steps - got]
        out.append((ks[0], ks[-1], sh, 0))
        got += len([K for K in range(ks[0], ks[-1] + 1) if will_search(K)])
        if got >= steps:
            break
    return out, got


def report_all(hosts, sols):
    for s in sorted(sols.tolist(), key=lambda s: (s[3], s[2], s[0])):
        h = hosts[s[3]] if s[3] >= 0 else hosts[0]
        if h.report(s[0], s[1], s[2], s[5], s[6]) != 0:
            raise RuntimeError("report failed")


def run_batched(ctx, steps, warmup):
    if warmup > 0:
        wu, _ = c4_sample(warmup, per_unit=1)
        ctx.search_batch(wu)
    units, nk = c4_sample(steps)
    tmp = tempfile.mkdtemp(prefix="ap26bench")
    hosts = []
    for i, (a, b, sh, _) in enumerate(units):
        d = os.path.join(tmp, str(i))
        os.makedirs(d)
        hosts.append(H.Host(a, b, sh, d, min_report=20))
    ctx.sync()
    t0 = time.perf_counter()
    sols, tot, st = ctx.search_batch(units)
    report_all(hosts, sols)
    ctx.sync()
    dt = time.perf_counter() - t0
    for h in hosts:
        h.close()
    return {"seconds": dt, "k": nk, "units": len(units), "stats": st.as_dict(),
            "shifts": sorted(set(u[2] for u in units))}


def run_plain(ctx, steps, warmup):
    ks = c3_ks(warmup + steps)
    tmp = tempfile.mkdtemp(prefix="ap26bench")
    h = H.Host(ks[0], ks[-1], SHIFT_C3, tmp, min_report=20)
    for K in ks[:warmup]:
        ctx.search_k(K, SHIFT_C3)
    ctx.sync()
    agg = {"sieve_ms": 0.0, "walk_ms": 0.0, "launches": 0, "positions": 0, "survivors": 0, "tests": 0}
    per_k = []
    t0 = time.perf_counter()
    for K in ks[warmup:]:
        tk = time.perf_counter()
        sols, st = ctx.search_k(K, SHIFT_C3)
        for s in sorted(sols.tolist(), key=lambda s: (s[2], s[0])):
            if h.report(s[0], s[1], s[2], s[5], s[6]) != 0:
                raise RuntimeError("report failed")
        per_k.append(time.perf_counter() - tk)
        agg["sieve_ms"] += st.ms_sieve
        agg["walk_ms"] += st.ms_walk
        agg["launches"] += st.sieve_launches
        agg["positions"] += st.positions
        agg["survivors"] += st.survivors
        agg["tests"] += st.prp_tests
    ctx.sync()
    dt = time.perf_counter() - t0
    h.close()
    n = len(ks) - warmup
    avg_launch_s = agg["sieve_ms"] / agg["launches"] / 1e3
    ops = agg["positions"] / agg["launches"] * LOOKUPS_PER_MASK_C3 * OPS_PER_LOOKUP
    return {"k_per_s": n / dt, "k": n, "ms_per_k": dt * 1e3 / n, "K_first": ks[warmup], "K_last": ks[-1],
            "sieve_ms_per_k": agg["sieve_ms"] / n, "walk_ms_per_k": agg["walk_ms"] / n,
            "survivors_per_k": agg["survivors"] / n, "prp_tests_per_k": agg["tests"] / n,
            "p95_k_latency_ms": float(np.percentile(per_k, 95) * 1e3),
            "sieve_roofline": {"achieved": ops / avg_launch_s / 1e12, "avg_launch_ms": avg_launch_s * 1e3,
                               "frac": ops / avg_launch_s / 1e12 / PEAK_INT32_TOPS}}


def run_prp(ctx, count=1 << 26, seed=26):
    rng = np.random.default_rng(seed)
    n = rng.integers(2 ** 58, 2 ** 63, size=count, dtype=np.uint64) | np.uint64(1)
    bl = np.floor(np.log2(n.astype(np.float64))).astype(np.int64) + 1
    squarings = int((bl - 2).sum())  # W(N) = bitlen(N) - 2, counted on the inputs
    d_n = ctx.dev_alloc(n.nbytes)
    d_o = ctx.dev_alloc(count)
    ctx.h2d(d_n, n)
    ctx.prp_batch_dev(d_n, d_o, count)
    times = [ctx.prp_batch_dev(d_n, d_o, count) for _ in range(5)]
    ctx.dev_free(d_n)
    ctx.dev_free(d_o)
    ms = float(np.median(times))
    tops = squarings * OPS_PER_SQUARING / (ms / 1e3) / 1e12
    return {"tests_per_s": count / (ms / 1e3), "ms": ms, "tests": count,
            "mean_squarings": squarings / count,
            "roofline": {"bound": "valu_int32", "achieved": tops, "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                         "frac": tops / PEAK_INT32_TOPS}}


def run_serving(ctx, kps, n_units=12, budget=3 << 30, seed=27, load=0.5):
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0, np.log(16), n_units))).astype(int)
    starts = rng.integers(1_000_000, 100_000_001, n_units)
    shifts = rng.choice([0, 640, 1280], n_units)
    units = [(int(s), int(s + L - 1), int(sh)) for s, L, sh in zip(starts, lens, shifts)]
    mean_k = np.mean([sum(will_search(K) for K in range(a, b + 1)) for a, b, _ in units])
    rate = load * kps / max(mean_k, 1e-9)   # units per second
    gaps = rng.exponential(1.0 / rate, n_units)
    arrive = np.cumsum(gaps) - gaps[0]
    lat = []
    with gpu.Server(ctx, max_units=64, mem_budget_bytes=budget) as srv:
        t0 = time.perf_counter()
        pending = list(range(n_units))
        tickets = {}
        while len(lat) < n_units:
            now = time.perf_counter() - t0
            while pending and arrive[pending[0]] <= now:
                i = pending.pop(0)
                tickets[i] = srv.submit(*units[i])
            if srv.stats()["active"] == 0 and pending:
                time.sleep(max(0.0, arrive[pending[0]] - (time.perf_counter() - t0)))
                continue
            srv.pump()
            for i, t in list(tickets.items()):
                r = srv.result(t)
                if r is not None:
                    lat.append(r[2])
                    del tickets[i]
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array(lat)
    return {"units": n_units, "rate_units_per_s": rate, "load": load, "k_per_s": st["k_done"] / wall,
            "k_per_round": st["k_per_round"], "p50_ms": float(np.percentile(lat, 50)),
            "p95_ms": float(np.percentile(lat, 95)), "p99_ms": float(np.percentile(lat, 99)),
            "unit_len": "floor(exp(U(0, ln 16)))"}


def cpu_baseline(threads, j1=2048):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = time.perf_counter()
    O.search_span(K0_C3, SHIFT_C3, 0, 10, 0, j1, threads=threads)
    dt = time.perf_counter() - t
    frac = j1 / 10840
    return {"value": frac / dt, "unit": "K/s", "cores": threads, "kind": "port",
            "sample": f"oracle search of K={K0_C3} SHIFT {SHIFT_C3}, all 10 shifts, n43 starts "
                      f"0..{j1 - 1} ({frac * 100:.1f}% of one K) in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-prp", action="store_true")
    ap.add_argument("--full-tune", action="store_true", help="autotune over all 27 candidates (9 variants x 3 work-item sizes)")
    ap.add_argument("--serve-loads", default="0.5",
                    help="offered loads of the serving leg, as fractions of the batched K/s (e.g. 0.25,0.5,1.0)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner: AP26_BATCH_VARIANT / AP26_BATCH_CHUNK or the defaults (0, 85)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the search shards by K with no inter-GPU traffic: each GPU would run its own units, "
              "and only the reports and per-unit hash totals would be gathered)", file=sys.stderr)
        sys.exit(2)

    ctx = gpu.Context(0)
    if args.no_tune:
        tune_best = {"variant": int(os.environ.get("AP26_BATCH_VARIANT", "0")),
                     "chunk": int(os.environ.get("AP26_BATCH_CHUNK", "85"))}
        tune_all = []
    else:
        tune_best, tune_all = ctx.autotune(K=K0_C3, shift=SHIFT_C3, quick=not args.full_tune)
    b = run_batched(ctx, args.steps, args.warmup)
    kps = b["k"] / b["seconds"]
    st = b["stats"]
    rounds = max(1, st["sieve_launches"])
    pos_per_launch = st["positions"] / rounds
    avg_launch_ms = st["ms_sieve"] / rounds
    ops = pos_per_launch * LOOKUPS_PER_MASK_MIX * OPS_PER_LOOKUP
    achieved = ops / (avg_launch_ms / 1e3) / 1e12 if avg_launch_ms > 0 else None
    k_per_launch = b["k"] / rounds
    alg_bytes = (st["survivors"] * SURVIVOR_BYTES + b["k"] * TABLE_BYTES_PER_K) / rounds
    pair = tune_best["variant"] == 0 or tune_best["variant"] >= 10
    kernel_name = ("sieve_persistent2_kernel" if pair else "sieve_persistent_kernel") + \
        f" (batched variant {tune_best['variant']}, chunk {tune_best['chunk']})"
    line = {
        "metric": METRIC,
        "value": kps,
        "unit": "K/s",
        "n_gpus": 1,
        "steps": b["k"],
        "warmup": args.warmup,
        "ms_per_step": b["seconds"] * 1e3 / b["k"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "batched persistent path: first min(len,4) K of the 256-unit mixed "
                               "workload (default_rng(26); len floor(exp(U(0,ln 4096))), start U[1e6,1e8], "
                               "SHIFT {0,640,1280}), one captured hipGraph per round of 16 K",
                   "k_searched": b["k"], "units": b["units"], "shifts": b["shifts"],
                   "rounds": rounds, "path": "batched",
                   "autotuned": {"variant": tune_best["variant"], "chunk": tune_best["chunk"],
                                 "candidates": len(tune_all)}},
        "roofline": {"bound": "valu_int32", "achieved": achieved, "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                     "frac": achieved / PEAK_INT32_TOPS if achieved else None,
                     "traffic": TRAFFIC_BYTES_PER_K * k_per_launch,
                     "hbm": {"algorithmic_bytes": alg_bytes, "traffic_bytes": TRAFFIC_BYTES_PER_K * k_per_launch,
                             "achieved_GBs": alg_bytes / (avg_launch_ms / 1e3) / 1e9,
                             "frac_of_8TBs": alg_bytes / (avg_launch_ms / 1e3) / 1e9 / PEAK_HBM_GBS},
                     "kernel": kernel_name, "avg_launch_ms": avg_launch_ms,
                     "ops_per_launch": ops,
                     "basis": f"{LOOKUPS_PER_MASK_MIX:.3f} lookups/mask x {OPS_PER_LOOKUP} int32 ops x "
                              f"{pos_per_launch:.4g} masks/launch (one launch = one round)"},
    }
    if not args.no_plain:
        p = run_plain(ctx, args.plain_steps, args.warmup)
        line["plain"] = p
        line["batched_over_plain"] = kps / p["k_per_s"]
    if not args.no_prp:
        line["prp"] = run_prp(ctx)
    if not args.no_serve:
        loads = [float(x) for x in args.serve_loads.split(",")]
        line["serving"] = run_serving(ctx, kps, load=loads[0])
        if len(loads) > 1:
            line["serving_sweep"] = [run_serving(ctx, kps, load=x) for x in loads]
    ctx.close()
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        line["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
