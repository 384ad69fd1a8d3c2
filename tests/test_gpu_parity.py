"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
inputs, and against the reference CPU build's golden files at full K."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from ap26_amd import gpu
from ap26_amd import host as H
from conftest import GOLDEN, read_sol

pytestmark = pytest.mark.gpu

MOD = 258559632607830
PRIM23 = 223092870


@pytest.mark.parametrize("K,shift", [(366384, 0), (366384, 576), (875859, 640 + 64 * 9),
                                     (976061, 1280), (44121555, 0), (99999989, 1280 + 64 * 5),
                                     (61 * 67 * 71 * 73, 64)])
def test_okok_tables(gpu_ctx, K, shift):
    assert np.array_equal(gpu_ctx.okok_table(K, shift), O.okok_table(K, shift))


def test_prp_small_exhaustive(gpu_ctx):
    n = np.arange(0, 1 << 17, dtype=np.uint64)
    assert np.array_equal(gpu_ctx.prp_batch(n), O.prp2_array(n))


def test_prp_edges(gpu_ctx):
    vals = [0, 1, 2, 3, 4, 5, 9, 2047, 3215031751, 3825123056546413051, 2 ** 61 - 1, 2 ** 62 - 57,
            2 ** 62 + 1, 2 ** 63 - 25, 2 ** 63 + 1, 2 ** 64 - 59, 2 ** 64 - 1, 2 ** 64 - 2]
    n = np.array(vals, dtype=np.uint64)
    assert gpu_ctx.prp_batch(n).tolist() == [int(O.prp2(v)) for v in vals]


def test_prp_microbench_full_against_oracle(gpu_ctx):
    """All 2^26 inputs of the PRP micro-benchmark (odd values uniform in
    [2^58, 2^63), numpy default_rng(26), as bench.py draws them): the device's
    booleans equal the oracle's SPRP on every one (threaded oracle)."""
    rng = np.random.default_rng(26)
    n = rng.integers(2 ** 58, 2 ** 63, size=1 << 26, dtype=np.uint64) | np.uint64(1)
    g = gpu_ctx.prp_batch(n)
    o = O.prp2_array(n, threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    assert np.array_equal(g, o) and 0 < o.sum() < len(o)


SPANS = [(366384, 0, 0, 10, 0, 3), (875859, 640, 0, 10, 100, 102), (48620635, 640, 3, 7, 5000, 5002),
         (44121555, 0, 0, 10, 10838, 10840), (976061, 0, 2, 3, 7000, 7004),
         (99999989, 1280, 9, 10, 1, 2), (1000003, 1280, 0, 10, 5419, 5421),
         (61 * 67 * 71 * 73, 0, 0, 10, 3000, 3002),
         (600000000, 16912, 0, 10, 4000, 4002)]  # largest K and SHIFT: terms just below 2^63


@pytest.mark.parametrize("variant", [0, 1, 2, 23, 31, 40, 42, 45, 46, 47, 48, 50, 52])
@pytest.mark.parametrize("span", SPANS)
def test_sieve_span_parity(variant, span):
    """Plain-path sieve variants (0 = 53, the rotated default with 512-byte
    rotation rows, seven compaction levels, queues 1.. in global memory and
    34 primes in LDS; 47 = the same with four levels;
    46 = 256-byte rows, the last queue in global memory, 37 primes in LDS;
    45 = every queue in LDS, the allocation-free entry point's; 48 = 47 with
    the last 4 level-0 primes read through L1 instead of LDS; 50 = 47 with
    the packed level-0 gathers on live lanes only, 52 = the same at 16 waves
    per CU; 31 = the
    shift-pair sieve without the rotation group, 40 = rotation group with
    level 0 on exactly its 16 primes, 42 = on 18 primes, 1/2 = the simple
    per-lane sieves, 23 = one-shift queued sieve) keep exactly the oracle's
    survivor set."""
    with gpu.Context(0, sieve_variant=variant) as ctx:
        g = np.sort(ctx.sieve_span(*span))
    o = np.sort(O.sieve_span(*span))
    assert len(o) > 0 and np.array_equal(g, o)


@pytest.mark.parametrize("span", [(366384, 0, 0, 10, 0, 20), (875862, 640, 0, 10, 2000, 2020),
                                  (44121555, 0, 0, 10, 10800, 10840), (61 * 67 * 71 * 73, 0, 0, 10, 100, 110),
                                  (600000000, 8000, 0, 10, 0, 20)])
def test_search_span_parity(gpu_ctx, span):
    g, st = gpu_ctx.search_span(*span)
    o, nsurv = O.search_span(*span, threads=4)
    assert st.survivors == nsurv
    assert sorted((int(a["k"]), int(a["first"])) for a in g) == \
        sorted((int(a["k"]), int(a["first"])) for a in o)


def _golden_units(golden_index):
    return [g for g in golden_index]


def test_full_k_against_reference_golden(gpu_ctx, golden_index, tmp_path):
    """Every golden unit (reference CPU build, min10): identical sorted SOL
    records and identical hash line."""
    assert len(golden_index) >= 12
    for g in golden_index:
        d = tmp_path / g["name"]
        d.mkdir()
        h = H.Host(g["kmin"], g["kmax"], g["shift"], str(d), min_report=10)
        for K in range(g["kmin"], g["kmax"] + 1):
            if not H.will_search(K):
                continue
            sols, st = gpu_ctx.search_k(K, g["shift"])
            for s in sorted(sols.tolist(), key=lambda s: (s[2], s[0])):
                # ReportSolution on the device's 9-base verification bits
                assert h.report(s[0], s[1], s[2], s[5], s[6]) == 0
        assert h.write_hash() == 0
        h.close()
        got = read_sol(str(d / "SOL-AP26.txt"))
        want = read_sol(os.path.join(GOLDEN, "ref_min10", g["name"] + ".sol"))
        assert got == want, g["name"]


def _oracle_masks(k, K, first):
    """Per-term bits from the oracle: bit j of the first word when term j is a
    base-2 strong PRP (ap26o_sprp), of the second when it is one to all nine
    bases 2..23 (ap26o_validate on the one-term progression, the reference's
    validate_ap26 rule, AP26.cpp:513-543)."""
    p2 = v9 = 0
    for j in range(min(k, 32)):
        n = first + j * K * PRIM23
        if n >= 3 and O.sprp(2, n):
            p2 |= 1 << j
            if O.validate(1, K, n, 9) == 1:
                v9 |= 1 << j
    return p2, v9


def test_device_verification_masks(gpu_ctx):
    """9-base verification on the device (validate_ap26 / val_base2_ap26,
    AP26.cpp:484-543) against the oracle's SPRP and validation: every
    report of two K, the reference's report vectors (APs holding base-2
    strong pseudoprimes), and APs built around A001262 pseudoprimes."""
    import json
    sols, _ = gpu_ctx.search_k(44121555, 0)
    rows = [(int(s["k"]), int(s["K"]), int(s["first"]), int(s["prp2_mask"]), int(s["valid_mask"])) for s in sols]
    sols2, _ = gpu_ctx.search_k(875859, 640)
    rows += [(int(s["k"]), int(s["K"]), int(s["first"]), int(s["prp2_mask"]), int(s["valid_mask"])) for s in sols2]
    assert len(rows) > 300
    for k, K, f, p2, v9 in rows:
        assert (p2, v9) == _oracle_masks(k, K, f), (k, K, f)
    vecs = json.load(open(os.path.join(GOLDEN, "report_vectors.json")))
    extra = [tuple(r) for v in vecs for r in v["reports"]]
    # pseudoprimes to base 2 that fail other bases, as first and middle terms
    for psp in (2047, 3277, 4033, 4681, 8321, 3215031751, 3825123056546413051):
        extra += [(12, 1, psp), (12, 3, psp - 5 * 3 * PRIM23) if psp > 5 * 3 * PRIM23 else (5, 1, psp)]
    arr = np.zeros(len(extra), dtype=gpu.SOL_DTYPE)
    for i, (k, K, f) in enumerate(extra):
        arr[i]["k"], arr[i]["K"], arr[i]["first"] = k, K, f
    out = gpu_ctx.verify_sols(arr)
    for i, (k, K, f) in enumerate(extra):
        assert (int(out[i]["prp2_mask"]), int(out[i]["valid_mask"])) == _oracle_masks(k, K, f), (k, K, f)


def test_survivor_count_of_reference_comment(gpu_ctx):
    """src/opencl/AP26.h:343: 38,838,420 sieve survivors for K=366384."""
    _, st = gpu_ctx.search_k(366384, 0)
    assert st.survivors == 38838420 and st.reports == 180


def test_full_k_survivor_invariants(gpu_ctx):
    """Size-independent properties of a full K's survivors: distinct, odd,
    < 2^62, free of 7..23, and no term n + i*STEP (i < 24) divisible by 61..541."""
    K, shift = 875859, 640
    surv = gpu_ctx.sieve_span(K, shift, 0, 10, 0, 10840, cap=1 << 26)
    assert 30_000_000 < len(surv) < 50_000_000
    assert len(np.unique(surv)) == len(surv)
    assert np.all(surv & 1 == 1) and surv.max() < 2 ** 62
    jj = surv // MOD
    assert jj.min() >= shift and jj.max() < shift + 640
    rng = np.random.default_rng(0)
    s = rng.choice(surv, 200_000, replace=False)
    for q in (7, 11, 13, 17, 19, 23):
        assert np.all(s % q != 0)
    step = K * PRIM23
    for p in [p for p in range(61, 542) if all(p % d for d in range(2, int(p ** .5) + 1))]:
        r, sp = s % p, step % p
        for i in range(24):
            assert np.all((r + i * sp) % p != 0)


def test_heavy_k_grows_survivor_buffer(gpu_ctx):
    """K = 61*67*71*73: four sieve primes divide STEP, so each rejects one
    residue instead of 24 and the sieve keeps about 5.4x the usual survivors,
    more than the initial 2^26-entry buffer.  The plain path grows its buffer,
    the batched path budgets the round for it, and both agree."""
    K = 61 * 67 * 71 * 73
    sols, st = gpu_ctx.search_k(K, 0)
    assert st.survivors > 1 << 26
    want = sorted((int(a["k"]), int(a["first"])) for a in sols)
    bs, tot, bst = gpu_ctx.search_batch([(366384, 366384, 0, 0), (K, K, 0, 0), (875859, 875859, 640, 0)])
    got = sorted((int(a["k"]), int(a["first"])) for a in bs if a["range_id"] == 1)
    assert got == want
    assert tot[1] == sum(f % 1000 + k for k, f in want)
    _, st2 = gpu_ctx.search_k(366384, 0)  # the grown buffer serves ordinary K as before
    assert st2.survivors == 38838420
    _, st3 = gpu_ctx.search_k(875859, 640)
    assert bst.survivors == st.survivors + st2.survivors + st3.survivors


def test_iter_start_is_a_suffix(gpu_ctx):
    full, _ = gpu_ctx.search_k(875862, 640)
    part, _ = gpu_ctx.search_k(875862, 640, iter_start=4)
    want = sorted((int(a["k"]), int(a["first"])) for a in full if a["shift_iter"] >= 4)
    assert sorted((int(a["k"]), int(a["first"])) for a in part) == want
    assert len(part) < len(full)


def test_edge_cases(gpu_ctx):
    g, st = gpu_ctx.search_span(366384, 0, 3, 3, 0, 10840)
    assert len(g) == 0 and st.survivors == 0
    assert len(gpu_ctx.sieve_span(366384, 0, 0, 10, 500, 500)) == 0
    lib = gpu.load()
    n = ctypes.c_int32()
    for args in [(0, 0, 0, 10, 0, 1), (366384, 0, 0, 11, 0, 1), (366384, 0, 0, 10, 0, 10841),
                 (366384, -64, 0, 10, 0, 1), (366384, 0, 5, 4, 0, 1),
                 (366384, 36000, 0, 10, 0, 1), (366384, 1 << 20, 0, 10, 0, 1),
                 (600000000, 20000, 0, 10, 0, 1),  # terms past 2^63
                 (600000000, 16912, 0, 10, 0, 1)]:  # first walked terms past 2^62
        rc = lib.ap26_search_span(gpu_ctx.h, *args, None, 0, ctypes.byref(n), None)
        assert rc == gpu.AP26_E_RANGE
    # capacity overflow reports the true count
    out = np.zeros(1, dtype=gpu.SOL_DTYPE)
    rc = lib.ap26_search_span(gpu_ctx.h, 366384, 0, 0, 10, 0, 20, out.ctypes.data_as(ctypes.c_void_p),
                              1, ctypes.byref(n), None)
    assert rc == gpu.AP26_E_OVERFLOW and n.value == 2
