"""Benchmark of the AP26 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-steps N]
                    [--serve-loads 0.25,0.5,1.0] [--no-cpu] [--no-serve]
                    [--no-plain] [--no-prp] [--no-tune] [--full-tune]

Prints ONE JSON line.  The headline `value` is K/sec (10-shift) of the batched
persistent path; a step is one searched K.  The timed K are a bounded sample
of the batched workload (256 units from numpy default_rng(26), length
floor(exp(U(0, ln 4096))), start uniform in [1e6, 1e8], SHIFT from
{0, 640, 1280}): units are visited in bit-reversed index order, round-robin
over the three SHIFTs, each giving its first 3 searched K as a sub-range
(ap26_amd/workloads.py: c4_sample), so any step count spreads over the whole
unit list, holds every SHIFT and has units running across round boundaries.
The whole unit is searched -- tables, all 10 shifts of the sieve, the SPRP
walk, AP check, per-range hash totals on the device -- and the host applies
ReportSolution (9-base check, hash, SOL records) to every report inside the
timed region.  After timing, every unit's SOL records and hash line are
compared with the reference CPU build's output for the same K
(tests/golden/ref_min10); a mismatch fails the run (exit status 1).

The batched engine's sieve variant and work-item size, then its walk's
compaction depth, are chosen first by the library's autotuner (ap26_autotune,
10 + 3 candidates timed on one K).

Beside it on the line:
  plain         the plain per-K path (one kernel sequence per K, host sync per
                K) on the first searched K from 48620635 at SHIFT 640, each K
                checked against the reference build's output;
  roofline      the batched path's persistent sieve kernel (the dominant one)
                priced on the INT32 VALU roofline from algorithmic counts,
                timed with HIP events on the library's stream, beside its VALU
                issue-slot use from rocprofv3 counters (a fixed factor per K);
  prp           the isolated SPRP micro-benchmark (2^26 odd values in
                [2^58, 2^63)) and its INT32 roofline, likewise;
  serving       continuous batching at a 3 GiB device budget: a saturated run
                (all units at t = 0) beside the batched engine on the same
                units with the same round size, then Poisson arrivals at 0.25,
                0.5 and 1.0 of the saturated rate; p50/p95/p99 unit latency;
                every unit's result compared with the batched run's (alone vs
                in the mix).  Unit length floor(exp(U(0, ln 16))): the
                workload's 1..8192 scaled down so a sweep finishes in the bench;
  cpu_baseline  the reference CPU build's own AVX2 search, compiled from its
                sources (oracle/_ref), one process per host core available
                to this process (at most 16), each on ~20 s of a K of the
                plain workload, half in the first and half in the last shift
                group; plus the
                oracle's SPRP on a sample of the micro-benchmark inputs,
                checked against the GPU's outputs.
"""
import argparse
import json
import os
import platform
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ap26_amd import gpu  # noqa: E402
from ap26_amd import host as H  # noqa: E402
from ap26_amd import workloads as W  # noqa: E402

METRIC = "K/sec (10-shift) + PRP tests/sec, batched vs plain GPU; p95 work-unit latency"
GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_min10")
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
CLOCK_HZ = 2.4e9
CUS = 256
PEAK_INT32_TOPS = CUS * 64 * CLOCK_HZ / 1e12  # 39.3 T int32 lane-ops/s (nominal)
PEAK_HBM_GBS = 8000.0
# Fixed factors from rocprofv3 counters (profiles/sieve_rot_counters.md,
# profiles/prp_kernel_valu.md): VALU lane-instructions (SQ_INSTS_VALU x 64)
# the sieve issues per searched K of the batched workload, and prp_kernel
# per modular squaring.  Divided by the launch time they give the VALU
# instruction rate; VALU_ISSUE_CEILING is the largest rate measured on this
# GPU by ap26_amd/bin/ceiling_probe (profiles/ceiling_probe.md), 1.256
# wave-instructions per CU-cycle (two-source 32-bit integer instructions at
# 32 waves per CU; three-source ones and multiplies reach 0.82-0.92), so
# rate / ceiling is the issue-slot use.  The nominal INT32
# peak above assumes 1.0.  Rates and ceiling are both wall-time rates at the
# nominal 2.4 GHz (the counters show the clock at 2.20-2.36 GHz under the
# sieve's load, box to box, so per-clock rates are 2-9 % higher).
SIEVE_VALU_LANE_INSTR_PER_K = 3.3361e12
PRP_VALU_LANE_INSTR_PER_SQUARING = 41.37
VALU_ISSUE_CEILING = 1.256
# HBM bytes per searched K moved by the persistent sieve kernel, from two
# rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over a 16-K round of this
# workload, each scaled by what the same counter reads for prp_kernel over
# 2^26 values, which reads 512 MiB and writes 64 MiB exactly
# (profiles/sieve_rot_traffic.md): (2 x 4,169,556 + 536,497,536) KiB / 16 K,
# the excess over the algorithmic bytes the per-wave global-memory queues of
# the deeper compaction levels, written back from L2 and read again.
TRAFFIC_BYTES_PER_K = 3.4869e10
# The batched engine's sieve when no variant is named (pick_pvariant(0) in
# ap26_amd/csrc/ap26_batch.hip, the table in ap26_pvariants_*.hip); used to
# name the kernel on the line.
DEFAULT_BATCH_VARIANT = 117
# Algorithmic bytes of the sieve: each survivor written once (8-byte value +
# 4-byte job index); the per-K tables (10 shifts x 23,783 words x 8 B) and
# n43 starts (10,840 x 8 B) read once.
SURVIVOR_BYTES = 12
TABLE_BYTES_PER_K = 10 * 23783 * 8 + 10840 * 8
SERVE_BUDGET = 3 << 30


# ---- reference fixtures ----------------------------------------------------

def read_sol(path):
    lines = open(path).read().splitlines()
    return sorted(tuple(int(x) for x in l.split()) for l in lines[:-1]), lines[-1]


def load_fixtures():
    """(K, SHIFT) -> fixture name of every single-K run of the reference build."""
    try:
        with open(os.path.join(GOLDEN, "index.json")) as f:
            idx = json.load(f)
    except OSError:
        return {}
    return {(g["kmin"], g["shift"]): g["name"] for g in idx if g["kmin"] == g["kmax"]}


def expected_unit(fix, kmin, kmax, shift, ks):
    """The reference build's SOL records and hash line for the unit
    kmin..kmax (searched K `ks`), composed from its single-K runs: the hash
    total of a unit is the sum of its K's totals (AP26.cpp:295-357).  None
    if a K has no fixture."""
    recs, T = [], 0
    for K in ks:
        name = fix.get((K, shift))
        if name is None:
            return None
        r, hl = read_sol(os.path.join(GOLDEN, name + ".sol"))
        recs += r
        T += int(hl[8:], 16)
    return sorted(recs), H.hash_line(kmin, kmax, H.cap_total(T))


def check_units(fix, runs):
    """runs: [(kmin, kmax, shift, searched K, SOL path)] -> parity summary."""
    out = {"units": len(runs), "verified": 0, "k_verified": 0, "unverified": [], "mismatches": []}
    for kmin, kmax, sh, ks, path in runs:
        want = expected_unit(fix, kmin, kmax, sh, ks)
        if want is None:
            out["unverified"].append([kmin, kmax, sh])
            continue
        got = read_sol(path)
        if got != want:
            out["mismatches"].append([kmin, kmax, sh])
        else:
            out["verified"] += 1
            out["k_verified"] += len(ks)
    return out


# ---- legs --------------------------------------------------------------------

def report_unit(h, sols):
    for s in sorted(sols.tolist(), key=lambda s: (s[2], s[0])):
        if h.report(s[0], s[1], s[2], s[5], s[6]) != 0:
            raise RuntimeError("report failed")


def run_batched(ctx, steps, warmup):
    if warmup > 0:
        order = W.c4_visit_order()
        units = W.c4_units()
        wu = []
        for i in reversed(order):  # units the sample reaches last
            ks = W.searched(units[i][0], units[i][1])[:1]
            if ks:
                wu.append((ks[0], ks[0], units[i][2], 0))
            if len(wu) == warmup:
                break
        ctx.search_batch(wu)
    units, idx, ks = W.c4_sample(steps)
    nk = sum(len(k) for k in ks)
    tmp = tempfile.mkdtemp(prefix="ap26bench")
    hosts, paths = [], []
    for i, (a, b, sh, _) in enumerate(units):
        d = os.path.join(tmp, str(i))
        os.makedirs(d)
        hosts.append(H.Host(a, b, sh, d, min_report=10))
        paths.append(os.path.join(d, "SOL-AP26.txt"))
    ctx.sync()
    t0 = time.perf_counter()
    sols, tot, st = ctx.search_batch(units)
    for i, h in enumerate(hosts):
        report_unit(h, sols[sols["range_id"] == i])
        if h.write_hash() != 0:
            raise RuntimeError("hash line")
    ctx.sync()
    dt = time.perf_counter() - t0
    for h in hosts:
        h.close()
    parity = check_units(load_fixtures(), [(a, b, sh, k, p) for (a, b, sh, _), k, p in zip(units, ks, paths)])
    crossing = 0
    pos = 0
    for k in ks:
        crossing += (pos // 16) != ((pos + len(k) - 1) // 16)
        pos += len(k)
    return {"seconds": dt, "k": nk, "units": len(units), "unit_index": idx, "stats": st.as_dict(),
            "shifts": sorted(set(u[2] for u in units)), "units_across_rounds": crossing, "parity": parity}


def run_plain(ctx, steps, warmup):
    ks = W.c3_ks(warmup + steps)
    for K in ks[:warmup]:
        ctx.search_k(K, W.C3_SHIFT)
    ctx.sync()
    tmp = tempfile.mkdtemp(prefix="ap26bench")
    agg = {"sieve_ms": 0.0, "walk_ms": 0.0, "launches": 0, "positions": 0, "survivors": 0, "tests": 0}
    per_k, runs = [], []
    t0 = time.perf_counter()
    for K in ks[warmup:]:
        tk = time.perf_counter()
        d = os.path.join(tmp, str(K))
        os.makedirs(d)
        h = H.Host(K, K, W.C3_SHIFT, d, min_report=10)
        sols, st = ctx.search_k(K, W.C3_SHIFT)
        report_unit(h, sols)
        if h.write_hash() != 0:
            raise RuntimeError("hash line")
        h.close()
        runs.append((K, K, W.C3_SHIFT, [K], os.path.join(d, "SOL-AP26.txt")))
        per_k.append(time.perf_counter() - tk)
        agg["sieve_ms"] += st.ms_sieve
        agg["walk_ms"] += st.ms_walk
        agg["launches"] += st.sieve_launches
        agg["positions"] += st.positions
        agg["survivors"] += st.survivors
        agg["tests"] += st.prp_tests
    ctx.sync()
    dt = time.perf_counter() - t0
    n = len(ks) - warmup
    avg_launch_s = agg["sieve_ms"] / agg["launches"] / 1e3
    ops = agg["positions"] / agg["launches"] * LOOKUPS_PER_MASK_C3 * OPS_PER_LOOKUP
    return {"k_per_s": n / dt, "k": n, "ms_per_k": dt * 1e3 / n, "K_first": ks[warmup], "K_last": ks[-1],
            "sieve_ms_per_k": agg["sieve_ms"] / n, "walk_ms_per_k": agg["walk_ms"] / n,
            "survivors_per_k": agg["survivors"] / n, "prp_tests_per_k": agg["tests"] / n,
            "p95_k_latency_ms": float(np.percentile(per_k, 95) * 1e3),
            "sieve_roofline": {"achieved": ops / avg_launch_s / 1e12, "avg_launch_ms": avg_launch_s * 1e3,
                               "frac": ops / avg_launch_s / 1e12 / PEAK_INT32_TOPS},
            "parity": check_units(load_fixtures(), runs)}


def prp_inputs(count=1 << 26, seed=26):
    rng = np.random.default_rng(seed)
    return rng.integers(2 ** 58, 2 ** 63, size=count, dtype=np.uint64) | np.uint64(1)


def run_prp(ctx, n):
    count = n.size
    bl = np.floor(np.log2(n.astype(np.float64))).astype(np.int64) + 1
    squarings = int((bl - 2).sum())  # W(N) = bitlen(N) - 2, counted on the inputs
    d_n = ctx.dev_alloc(n.nbytes)
    d_o = ctx.dev_alloc(count)
    ctx.h2d(d_n, n)
    ctx.prp_batch_dev(d_n, d_o, count)
    times = [ctx.prp_batch_dev(d_n, d_o, count) for _ in range(5)]
    out = np.zeros(count, dtype=np.uint8)
    ctx.d2h(out, d_o)
    ctx.dev_free(d_n)
    ctx.dev_free(d_o)
    ms = float(np.median(times))
    tops = squarings * OPS_PER_SQUARING / (ms / 1e3) / 1e12
    rate = PRP_VALU_LANE_INSTR_PER_SQUARING * squarings / 64 / (ms / 1e3 * CLOCK_HZ * CUS)
    return out, {"tests_per_s": count / (ms / 1e3), "ms": ms, "tests": count,
                 "mean_squarings": squarings / count, "prp_count": int(out.sum()),
                 "roofline": {"bound": "valu_int32", "achieved": tops, "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                              "frac": tops / PEAK_INT32_TOPS,
                              "valu_instr_per_cu_cycle": rate, "valu_issue_frac": rate / VALU_ISSUE_CEILING,
                              "basis": f"{OPS_PER_SQUARING} int32 ops per squaring x bitlen(N)-2 squarings; "
                                       f"issue: {PRP_VALU_LANE_INSTR_PER_SQUARING} VALU lane-instr per squaring "
                                       f"(rocprofv3 SQ_INSTS_VALU), ceiling {VALU_ISSUE_CEILING} wave-instr "
                                       "per CU-cycle at 2.4 GHz"}}


def server_run(ctx, units, arrive, budget):
    """Submit unit i at arrive[i] s (None: all at once), pump until all are
    done; per-unit results and latencies, K/s over the wall time."""
    res = {}
    with gpu.Server(ctx, max_units=max(64, len(units)), mem_budget_bytes=budget) as srv:
        t0 = time.perf_counter()
        pending = list(range(len(units)))
        tickets = {}
        while len(res) < len(units):
            now = time.perf_counter() - t0
            while pending and (arrive is None or arrive[pending[0]] <= now):
                i = pending.pop(0)
                tickets[i] = srv.submit(*units[i])
            if srv.stats()["active"] == 0 and pending:
                time.sleep(max(0.0, arrive[pending[0]] - (time.perf_counter() - t0)))
                continue
            srv.pump()
            for i, t in list(tickets.items()):
                r = srv.result(t)
                if r is not None:
                    res[i] = r
                    del tickets[i]
        wall = time.perf_counter() - t0
        st = srv.stats()
    return res, wall, st


def key(sols):
    return sorted((int(x["k"]), int(x["K"]), int(x["first"])) for x in sols)


def run_serving(ctx, loads, n_units=12, seed=27, budget=SERVE_BUDGET):
    units = W.c5_units(n_units, seed=seed, max_len=16)
    nks = [len(W.searched(a, b)) for a, b, _ in units]
    total_k = sum(nks)
    # saturated: every unit at t = 0
    res0, wall0, st0 = server_run(ctx, units, None, budget)
    R = st0["k_per_round"]
    kps_sat = total_k / wall0
    # the batched engine on the same units with the same round size (same
    # device budget), in its own context
    os.environ["AP26_BATCH_ROUND"] = str(R)
    with gpu.Context(0) as bctx:
        bctx.search_batch([(units[0][0], units[0][0], units[0][2], 0)])
        t = time.perf_counter()
        sols, tot, _ = bctx.search_batch([(a, b, sh, 0) for a, b, sh in units])
        kps_batch = total_k / (time.perf_counter() - t)
    del os.environ["AP26_BATCH_ROUND"]
    alone = [(key(sols[sols["range_id"] == i]), int(tot[i])) for i in range(len(units))]
    mismatches = sum((key(res0[i][0]), res0[i][1]) != alone[i] for i in range(len(units)))
    out = {"units": n_units, "searched_k": total_k, "k_per_round": R, "budget_bytes": budget,
           "unit_len": "floor(exp(U(0, ln 16))) (the workload's ln 8192 scaled down)",
           "saturated": {"k_per_s": kps_sat, "wall_s": wall0,
                         "p50_ms": float(np.percentile([res0[i][2] for i in res0], 50)),
                         "p95_ms": float(np.percentile([res0[i][2] for i in res0], 95))},
           "batched_same_round_k_per_s": kps_batch, "saturated_over_batched": kps_sat / kps_batch,
           "sweep": []}
    rng = np.random.default_rng(seed + 1)
    for load in loads:
        rate = load * kps_sat / (total_k / n_units)  # units per second
        gaps = rng.exponential(1.0 / rate, n_units)
        arrive = np.cumsum(gaps) - gaps[0]
        res, wall, st = server_run(ctx, units, arrive, budget)
        lat = np.array([res[i][2] for i in range(n_units)])
        mismatches += sum((key(res[i][0]), res[i][1]) != alone[i] for i in range(n_units))
        progress(f"serving load {load}: p95 {np.percentile(lat, 95):.0f} ms")
        out["sweep"].append({"load": load, "rate_units_per_s": rate, "k_per_s": total_k / wall,
                             "p50_ms": float(np.percentile(lat, 50)), "p95_ms": float(np.percentile(lat, 95)),
                             "p99_ms": float(np.percentile(lat, 99))})
    out["parity"] = {"runs": 1 + len(loads), "unit_mismatches_vs_batched_alone": int(mismatches)}
    return out


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


REF_SAMPLE = os.path.join(ROOT, "oracle", "_ref", "ref_cpu_sample")


def n43_per_block():
    """n43 starts in each of the 85 progress blocks of one shift group: the
    reference's (i31, i37) blocks and, inside each, its i41 x i3 x i5 loops
    with their guards (src/cpu/cpuavx2.cpp:377-393); they sum to 10,840."""
    counts = []
    for i31 in range(7):
        for i37 in range(13):
            if i37 - i31 <= 10 and i31 - i37 <= 4:
                n41 = sum(1 for i41 in range(17)
                          if i41 - i31 <= 14 and i41 - i37 <= 14 and i31 - i41 <= 4 and i37 - i41 <= 10)
                counts.append(8 * n41)
    return counts


def reference_cpu_sample(procs, seconds=20.0):
    """The reference CPU build's own AVX2 search (Search_avx2, compiled from
    the reference's sources into oracle/_ref) on a bounded sample: one
    process per core, each on a different K of the plain workload at SHIFT
    640, stopped after `seconds` at a boundary of its n43 blocks
    (oracle/ref_cpu_sample.cpp).  The search runs the K's ten shifts as three
    groups of four vector lanes (shifts 0-3, 4-7, and 8-9 with two lanes
    masked), so half of the processes time the first group and half the
    last; the blocks done are counted in n43 starts (n43_per_block), and one
    K costs two first-group passes plus one last-group pass over the 10,840
    starts.  Returns (K list, seconds per K per core, detail)."""
    import subprocess
    tmp = tempfile.mkdtemp(prefix="ap26cpu")
    ks = W.c3_ks(procs)
    iters = [0 if i % 2 == 0 else 8 for i in range(len(ks))]
    per_block = np.cumsum([0] + n43_per_block())
    runs = []
    for i, (K, it) in enumerate(zip(ks, iters)):
        d = os.path.join(tmp, str(i))
        os.makedirs(d)
        runs.append(subprocess.Popen([REF_SAMPLE, str(K), str(W.C3_SHIFT), str(seconds), str(it)], cwd=d,
                                     stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True))
    rates = {0: [], 8: []}
    for p, it in zip(runs, iters):
        out, _ = p.communicate(timeout=seconds * 4 + 120)
        f = out.split()
        blocks, secs = int(f[f.index("blocks") + 1]), float(f[f.index("seconds") + 1])
        if p.returncode != 0 or blocks < 1 or secs <= 0:
            raise RuntimeError("reference CPU sample failed")
        rates[it].append(float(per_block[blocks]) / secs)  # n43 starts per second in this group
    if not rates[8]:  # one core: time the first group only and price the last group like it
        rates[8] = rates[0]
    g1, g3 = float(np.mean(rates[0])), float(np.mean(rates[8]))
    s_per_k = 2 * 10840 / g1 + 10840 / g3
    return ks, s_per_k, {"first_group_n43_per_s": g1, "last_group_n43_per_s": g3,
                         "last_over_first_cost": g1 / g3}


def cpu_baseline(n_prp, gpu_prp, j1=2048):
    """The CPU baseline on this machine's host cores: the reference CPU build
    itself (kind "reference") when oracle/_ref holds its sampler, else the
    oracle's scalar restatement (kind "port") on a slice of one K; plus the
    oracle's SPRP on 2^16 micro-benchmark inputs, checked against the GPU's
    outputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    model, avail = cpu_info()
    threads = max(1, min(16, avail))
    s = n_prp[: 1 << 16]
    t = time.perf_counter()
    want = O.prp2_array(s)
    dtp = time.perf_counter() - t
    prp = {"tests_per_s_1thread": s.size / dtp, "sample": int(s.size),
           "mismatches_vs_gpu": int((want != gpu_prp[: s.size]).sum()) if gpu_prp is not None else None}
    if os.access(REF_SAMPLE, os.X_OK):
        ks, s_per_k, detail = reference_cpu_sample(threads)
        return {"value": threads / s_per_k, "unit": "K/s", "cores": threads, "kind": "reference",
                "cpu_model": model, "cores_available": avail, "s_per_k_per_core": s_per_k,
                "sample": f"the reference CPU build's AVX2 search (Search_avx2, built from its sources), "
                          f"{threads} processes, one per core, each ~20 s of one K (K={ks[0]}..{ks[-1]}, "
                          f"SHIFT {W.C3_SHIFT}), half in the first shift group (shifts 0-3), half in the "
                          f"last (shifts 8-9), timed at its n43-block progress points; one K = two "
                          f"first-group passes + one last-group pass",
                "detail": detail, "prp": prp}
    t = time.perf_counter()
    O.search_span(W.C3_K0, W.C3_SHIFT, 0, 10, 0, j1, threads=threads)
    dt = time.perf_counter() - t
    frac = j1 / 10840
    return {"value": frac / dt, "unit": "K/s", "cores": threads, "kind": "port",
            "cpu_model": model, "cores_available": avail, "s_per_k_per_core": dt / frac * threads,
            "sample": f"oracle search of K={W.C3_K0} SHIFT {W.C3_SHIFT}, all 10 shifts, n43 starts "
                      f"0..{j1 - 1} ({frac * 100:.1f}% of one K) in {dt:.1f} s on {threads} threads",
            "prp": prp}


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


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
    ap.add_argument("--full-tune", action="store_true", help="autotune over all 51 candidates (12 variants x 4 work-item sizes, then 3 walk depths)")
    ap.add_argument("--serve-loads", default="0.25,0.5,1.0",
                    help="offered loads of the serving sweep, as fractions of the saturated serving K/s")
    ap.add_argument("--serve-units", type=int, default=10)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner: AP26_BATCH_VARIANT / AP26_BATCH_CHUNK or the defaults (0, 11)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the search shards by K with no inter-GPU traffic: each GPU would run its own units, "
              "and only the reports and per-unit hash totals would be gathered)", file=sys.stderr)
        sys.exit(2)

    ctx = gpu.Context(0)
    if args.no_tune:
        tune_best = {"variant": int(os.environ.get("AP26_BATCH_VARIANT", "0")),
                     "chunk": int(os.environ.get("AP26_BATCH_CHUNK", "11")),
                     "walk_depth": int(os.environ.get("AP26_WALK_DEPTH", "1"))}
        tune_all = []
    else:
        tune_best, tune_all = ctx.autotune(K=W.C3_K0 + 2, shift=W.C3_SHIFT, quick=not args.full_tune)
    progress(f"autotuned: variant {tune_best['variant']}, chunk {tune_best['chunk']}, "
             f"walk depth {tune_best['walk_depth']}")
    b = run_batched(ctx, args.steps, args.warmup)
    kps = b["k"] / b["seconds"]
    progress(f"batched: {b['k']} K in {b['seconds']:.2f} s, parity {b['parity']['verified']}/{b['units']} units")
    st = b["stats"]
    rounds = max(1, st["sieve_launches"])
    pos_per_launch = st["positions"] / rounds
    avg_launch_ms = st["ms_sieve"] / rounds
    ops = pos_per_launch * LOOKUPS_PER_MASK_MIX * OPS_PER_LOOKUP
    achieved = ops / (avg_launch_ms / 1e3) / 1e12 if avg_launch_ms > 0 else None
    k_per_launch = b["k"] / rounds
    alg_bytes = (st["survivors"] * SURVIVOR_BYTES + b["k"] * TABLE_BYTES_PER_K) / rounds
    rate = SIEVE_VALU_LANE_INSTR_PER_K * k_per_launch / 64 / (avg_launch_ms / 1e3 * CLOCK_HZ * CUS) \
        if avg_launch_ms else None
    variant = tune_best["variant"] or DEFAULT_BATCH_VARIANT
    kernel_name = f"sieve_persistent2_kernel (batched variant {variant}, chunk {tune_best['chunk']})"
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
        "config": {"workload": "batched persistent path: a strided sample of the 256-unit mixed workload "
                               "(default_rng(26); len floor(exp(U(0,ln 4096))), start U[1e6,1e8], "
                               "SHIFT {0,640,1280}; units in bit-reversed order round-robin over SHIFT, "
                               "first 3 searched K of each), one captured hipGraph pair per round of 16 K",
                   "k_searched": b["k"], "units": b["units"], "unit_index": b["unit_index"], "shifts": b["shifts"],
                   "units_across_rounds": b["units_across_rounds"], "rounds": rounds, "path": "batched",
                   "autotuned": {"variant": tune_best["variant"], "chunk": tune_best["chunk"],
                                 "walk_depth": tune_best["walk_depth"], "candidates": len(tune_all)}},
        "parity": b["parity"],
        "roofline": {"bound": "valu_int32", "achieved": achieved, "peak": PEAK_INT32_TOPS, "unit": "TOP/s",
                     "frac": achieved / PEAK_INT32_TOPS if achieved else None,
                     "valu_instr_per_cu_cycle": rate,
                     "valu_issue_frac": rate / VALU_ISSUE_CEILING if rate else None,
                     "traffic": TRAFFIC_BYTES_PER_K * k_per_launch,
                     "hbm": {"algorithmic_bytes": alg_bytes, "traffic_bytes": TRAFFIC_BYTES_PER_K * k_per_launch,
                             "achieved_GBs": alg_bytes / (avg_launch_ms / 1e3) / 1e9,
                             "frac_of_8TBs": alg_bytes / (avg_launch_ms / 1e3) / 1e9 / PEAK_HBM_GBS},
                     "kernel": kernel_name, "avg_launch_ms": avg_launch_ms,
                     "ops_per_launch": ops,
                     "basis": f"{LOOKUPS_PER_MASK_MIX:.3f} lookups/mask x {OPS_PER_LOOKUP} int32 ops x "
                              f"{pos_per_launch:.4g} masks/launch (one launch = one round); issue: "
                              f"{SIEVE_VALU_LANE_INSTR_PER_K:.4g} VALU lane-instr per K (rocprofv3 "
                              f"SQ_INSTS_VALU), ceiling {VALU_ISSUE_CEILING} wave-instr per CU-cycle at 2.4 GHz"},
    }
    if not args.no_plain:
        p = run_plain(ctx, args.plain_steps, args.warmup)
        line["plain"] = p
        line["batched_over_plain"] = kps / p["k_per_s"]
        progress(f"plain: {p['k_per_s']:.3f} K/s")
    gpu_prp, n_prp = None, None
    if not args.no_prp or not args.no_cpu:
        n_prp = prp_inputs()
    if not args.no_prp:
        gpu_prp, line["prp"] = run_prp(ctx, n_prp)
        progress(f"prp: {line['prp']['tests_per_s']:.4g} tests/s")
    if not args.no_serve:
        loads = [float(x) for x in args.serve_loads.split(",") if x]
        line["serving"] = run_serving(ctx, loads, n_units=args.serve_units)
        progress("serving done")
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n_prp, gpu_prp)
    print(json.dumps(line), flush=True)
    bad = line["parity"]["mismatches"] or ("plain" in line and line["plain"]["parity"]["mismatches"]) or \
        ("serving" in line and line["serving"]["parity"]["unit_mismatches_vs_batched_alone"]) or \
        ("cpu_baseline" in line and line["cpu_baseline"]["prp"]["mismatches_vs_gpu"])
    if bad:
        print("bench.py: results differ from the reference's", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
