"""Batched engine (ap26_search_batch) and serving loop (ap26_server_*):
every range / unit gets the same reports and hash total as the plain per-K
path and the reference golden files, alone and in a mix."""
import os

import numpy as np
import pytest

from ap26_amd import gpu
from ap26_amd import host as H
from ap26_amd import workloads as W
from conftest import GOLDEN, read_sol

pytestmark = pytest.mark.gpu


def will_search(K):
    return all(K % p for p in (29, 31, 37, 41, 43, 47, 53, 59))


def plain_reports(ctx, kmin, kmax, shift, it=0):
    out = []
    for K in range(kmin, kmax + 1):
        if will_search(K):
            s, _ = ctx.search_k(K, shift, it if K == kmin else 0)
            out += [(int(x["k"]), int(x["K"]), int(x["first"])) for x in s]
    return sorted(out)


def key(sols):
    return sorted((int(x["k"]), int(x["K"]), int(x["first"])) for x in sols)


RANGES = [(366384, 366384, 0, 0), (875858, 875860, 640, 0), (44121555, 44121555, 0, 0),
          (875861, 875861, 640, 4), (99999989, 99999989, 1280, 0), (976061, 976062, 0, 0),
          (29, 29, 0, 0), (1000003, 1000003, 1280, 9)]


def test_batch_matches_plain_path_and_replays(gpu_ctx):
    want = [plain_reports(gpu_ctx, *r) for r in RANGES]
    for _ in range(2):  # the second call replays the captured graph
        sols, tot, st = gpu_ctx.search_batch(RANGES)
        assert st.k_searched == sum(1 for a, b, _, _ in RANGES for K in range(a, b + 1) if will_search(K))
        for i in range(len(RANGES)):
            got = key([x for x in sols if x["range_id"] == i])
            assert got == want[i], i
            assert tot[i] == sum(f % 1000 + k for k, _, f in got)


def test_batch_many_rounds_against_golden(gpu_ctx, golden_index):
    """More K than one round holds (16): the graph is replayed back to back;
    per-unit hash lines equal the reference's."""
    units = [g for g in golden_index if g["kmin"] == g["kmax"]]
    assert len(units) >= 12
    units = units + units[:6]  # > 16 K: at least two rounds
    ranges = [(g["kmin"], g["kmax"], g["shift"], 0) for g in units]
    sols, tot, st = gpu_ctx.search_batch(ranges)
    for i, g in enumerate(units):
        recs, hline = read_sol(os.path.join(GOLDEN, "ref_min10", g["name"] + ".sol"))
        assert H.hash_line(g["kmin"], g["kmax"], H.cap_total(int(tot[i]))) == hline, g["name"]
        got = [r for r in key([x for x in sols if x["range_id"] == i]) if r[0] >= 10]
        assert got == recs


def test_batch_segments_equal_one_segment(gpu_ctx, golden_index, monkeypatch):
    """A call runs its K in segments (default 2048 searched K), each with its
    own verification and report copy, and halves a segment whose reports
    exceed the budget: reports, per-range totals and counts equal a
    one-segment call.  Here segments of 5 K and a budget of 400 reports
    (about 2 K) force both the split and the halving."""
    units = [g for g in golden_index if g["kmin"] == g["kmax"]][:12]
    ranges = [(g["kmin"], g["kmax"], g["shift"], 0) for g in units]
    ranges.append((875859, 875862, 640, 2))  # a multi-K range resumed at ITER 2
    sols0, tot0, st0 = gpu_ctx.search_batch(ranges)
    monkeypatch.setenv("AP26_BATCH_SEG_K", "5")
    monkeypatch.setenv("AP26_BATCH_SOL_LIMIT", "400")
    sols1, tot1, st1 = gpu_ctx.search_batch(ranges)
    for i in range(len(ranges)):
        assert key([x for x in sols1 if x["range_id"] == i]) == key([x for x in sols0 if x["range_id"] == i]), i
    assert list(tot1) == list(tot0)
    d0, d1 = st0.as_dict(), st1.as_dict()
    for f in ("k_searched", "positions", "survivors", "prp_tests", "reports"):
        assert d1[f] == d0[f], f


@pytest.mark.parametrize("variant", ["0", "95", "126", "132", "138", "85", "75", "76", "98", "92", "103", "105", "107",
                                     "70", "54", "52", "56", "67", "28", "10", "9", "12", "13", "32"])
def test_batch_sieve_matches_plain_survivors(variant, monkeypatch):
    """The batched sieve (shift-pair kernels with and without the bank-rotated
    level-0 group -- 0 = 117 (seven compaction levels), 126 (seven, other
    splits), 132 (117 with non-temporal deep queues), 138 (seven levels with
    the level-1 queue in LDS), 95 (four), 98, 92 with 512-byte rotation rows and deeper
    queues in global memory, 103 = 95 with the last 4 level-0 primes read
    through L1, 105 = 95 with the packed level-0 gathers on live lanes
    only, 107 = the same at 16 waves per CU, 85, 75, 76 with deeper queues in global memory
    and 256-byte rows, 70, 54, 52, 56, 67 with
    every queue in LDS -- and one-shift persistent
    kernels) keeps exactly the plain path's survivors, including the last,
    partly filled stride of n53 blocks of every work item and a resume at an
    odd ITER (half of a shift pair skipped)."""
    monkeypatch.setenv("AP26_BATCH_VARIANT", variant)
    with gpu.Context(0) as ctx:
        for K, sh, it in [(366384, 0, 0), (875859, 640, 3), (99999989, 1280, 9)]:
            want = np.sort(ctx.sieve_span(K, sh, it, 10, 0, 10840, cap=1 << 26))
            got = np.sort(ctx.batch_sieve_k(K, sh, it))
            assert len(got) == len(want) and np.array_equal(got, want), (K, sh, it)


def test_batch_edge_cases(gpu_ctx):
    sols, tot, st = gpu_ctx.search_batch([(29, 31, 0, 0)])  # 29, 31 skipped, 30 searched
    assert st.k_searched == 1 and st.k_skipped == 2
    sols, tot, st = gpu_ctx.search_batch([(366384, 366384, 0, 10)])  # resumed past the last shift
    assert len(sols) == 0 and tot[0] == 0
    with pytest.raises(gpu.Ap26Error):
        gpu_ctx.search_batch([(5, 4, 0, 0), (0, 3, 0, 0)])
    with pytest.raises(gpu.Ap26Error):  # terms past 2^63
        gpu_ctx.search_batch([(366384, 366384, 36000, 0)])
    # the largest K at a large SHIFT whose walks stay below 2^62: batched
    # equals plain, survivors and reports
    sols, tot, st = gpu_ctx.search_batch([(600000000, 600000000, 8000, 0)])
    psols, pst = gpu_ctx.search_k(600000000, 8000)
    assert st.survivors == pst.survivors > 30_000_000 and st.walk_past_limit == pst.walk_past_limit == 0
    assert key(sols) == key(psols)


def test_walk_limit_of_the_reference(gpu_ctx):
    """The reference CPU build's PrimeQ exits on a walked term >= 2^62
    (src/common/PrimeQ_x86.h:160-166).  K = 6e8 at SHIFT 16912: the first
    walked term of the largest candidates is past 2^62, so the search is
    refused before any GPU work.  At SHIFT 14000 it is not, but walks reach
    2^62 at their third term: both paths run the K, count the same walked
    terms past 2^62 and fail with AP26_E_RANGE instead of reporting."""
    for call in (lambda: gpu_ctx.search_k(600000000, 16912),
                 lambda: gpu_ctx.search_batch([(600000000, 600000000, 16912, 0)])):
        with pytest.raises(gpu.Ap26Error) as e:
            call()
        assert e.value.code == gpu.AP26_E_RANGE
    with pytest.raises(gpu.Ap26Error) as ep:
        gpu_ctx.search_k(600000000, 14000)
    with pytest.raises(gpu.Ap26Error) as eb:
        gpu_ctx.search_batch([(600000000, 600000000, 14000, 0), (366384, 366384, 0, 0)])
    assert ep.value.code == eb.value.code == gpu.AP26_E_RANGE
    assert ep.value.stats.walk_past_limit == eb.value.stats.walk_past_limit > 0
    # the same unit in the server fails alone; its neighbour completes
    with gpu.Server(gpu_ctx, max_units=4, mem_budget_bytes=3 << 30) as srv:
        bad = srv.submit(600000000, 600000000, 14000)
        good = srv.submit(366384, 366384, 0)
        with pytest.raises(gpu.Ap26Error):  # refused at admission
            srv.submit(600000000, 600000000, 16912)
        res, err = None, None
        while res is None or err is None:
            srv.pump()
            if res is None:
                res = srv.result(good)
            if err is None:
                try:
                    if srv.result(bad) is not None:
                        err = "completed"
                except gpu.Ap26Error as x:
                    err = x.code
        assert err == gpu.AP26_E_RANGE
        want = plain_reports(gpu_ctx, 366384, 366384, 0)
        assert key(res[0]) == want and res[1] == sum(f % 1000 + k for k, _, f in want)


def test_server_alone_and_in_mix(gpu_ctx):
    units = [(875859, 875860, 640, 0), (366384, 366384, 0, 0), (875861, 875863, 640, 0),
             (44121555, 44121555, 0, 0), (1000003, 1000003, 1280, 3)]
    alone = []
    with gpu.Server(gpu_ctx, max_units=8, mem_budget_bytes=3 << 30) as srv:
        for u in units:
            t = srv.submit(*u)
            r = None
            while r is None:
                srv.pump()
                r = srv.result(t)
            alone.append((key(r[0]), r[1]))
    with gpu.Server(gpu_ctx, max_units=8, mem_budget_bytes=3 << 30) as srv:
        assert srv.stats()["k_per_round"] == 5  # 3 GiB / ~0.61 GB per in-flight K
        tickets = [srv.submit(*u) for u in units[:3]]
        srv.pump()                      # the first round runs with three units
        tickets += [srv.submit(*u) for u in units[3:]]  # late arrivals join
        res = {}
        while len(res) < len(units):
            srv.pump()
            for i, t in enumerate(tickets):
                if i not in res:
                    r = srv.result(t)
                    if r is not None:
                        res[i] = r
        for i in range(len(units)):
            assert (key(res[i][0]), res[i][1]) == alone[i], i
            assert res[i][2] > 0
        st = srv.stats()
        assert st["units_done"] == len(units) and st["active"] == 0
    # the plain path agrees
    for i, u in enumerate(units):
        assert alone[i][0] == plain_reports(gpu_ctx, *u)


def test_server_heavy_k(gpu_ctx):
    """A K with 5.4x the usual survivors (61*67*71*73 divides it) fills more
    than a 1 GiB server's round buffers on its own: the server runs it in a
    round of its own, grown to fit, and the unit beside it still matches."""
    K = 61 * 67 * 71 * 73
    units = [(K, K, 0), (875859, 875860, 640)]
    with gpu.Server(gpu_ctx, max_units=4, mem_budget_bytes=1 << 30) as srv:
        assert srv.stats()["k_per_round"] == 1
        tickets = [srv.submit(*u) for u in units]
        res = {}
        while len(res) < len(units):
            srv.pump()
            for i, t in enumerate(tickets):
                if i not in res:
                    r = srv.result(t)
                    if r is not None:
                        res[i] = r
    for i, (a, b, sh) in enumerate(units):
        want = plain_reports(gpu_ctx, a, b, sh)
        assert key(res[i][0]) == want, i
        assert res[i][1] == sum(f % 1000 + k for k, _, f in want)


def test_server_progress(gpu_ctx):
    """Per-unit progress while a unit runs: searched K done / in the unit."""
    with gpu.Server(gpu_ctx, max_units=4, mem_budget_bytes=1 << 30) as srv:  # one K per round
        t = srv.submit(875858, 875862, 640)
        assert srv.progress(t) == (0.0, 0, 4)
        seen, r = [], None
        while r is None:
            srv.pump()
            r = srv.result(t)
            if r is None:
                seen.append(srv.progress(t))
        assert [s[1] for s in seen] == [1, 2, 3] and seen[0][0] == 0.25
        with pytest.raises(gpu.Ap26Error):
            srv.progress(t)  # collected


def test_server_admission_limit(gpu_ctx):
    with gpu.Server(gpu_ctx, max_units=2, mem_budget_bytes=1 << 30) as srv:
        a = srv.submit(366384, 366384, 0)
        b = srv.submit(366385, 366385, 0)
        assert a is not None and b is not None
        assert srv.submit(366386, 366386, 0) is None  # full until a unit finishes
        done = set()
        while len(done) < 2:
            srv.pump()
            for t in (a, b):
                if t not in done and srv.result(t) is not None:
                    done.add(t)
        assert srv.submit(366386, 366386, 0) is not None
        e = srv.submit(29, 29, 0)  # nothing to search: finished at once
        assert srv.result(e) is not None


def test_autotune_picks_and_caches(tmp_path, monkeypatch):
    """Stage 1 picks the sieve (variant, chunk) by device time at walk depth
    1; stage 2 times the walk depths 1, 2, 4 with that sieve and picks by the
    walk's time; the cache carries all three and a new context adopts it."""
    monkeypatch.delenv("AP26_WALK_DEPTH", raising=False)
    cache = tmp_path / "tune.txt"
    with gpu.Context(0) as ctx:
        best, tried = ctx.autotune(K=875859, shift=640, quick=True, cache_path=str(cache))
        assert len(tried) == 13
        sieve, walk = tried[:10], tried[10:]
        assert all(t["walk_depth"] == 1 for t in sieve)
        first = min(sieve, key=lambda t: t["ms_per_k"])
        assert (best["variant"], best["chunk"]) == (first["variant"], first["chunk"])
        assert [t["walk_depth"] for t in walk] == [1, 2, 4]
        assert all((t["variant"], t["chunk"]) == (best["variant"], best["chunk"]) for t in walk)
        assert best["ms_walk"] == min(t["ms_walk"] for t in walk)
        assert 0 < best["ms_walk"] < best["ms_per_k"]
        want = plain_reports(ctx, 875859, 875859, 640)
        sols, tot, st = ctx.search_batch([(875859, 875859, 640, 0)])  # built with the choice
        assert key(sols) == want
    fields = cache.read_text().split()
    assert fields[:2] == ["ap26-tune", "2"]
    assert [int(x) for x in fields[3:6]] == [best["variant"], best["chunk"], best["walk_depth"]]
    monkeypatch.setenv("AP26_TUNE_CACHE", str(cache))
    with gpu.Context(0) as ctx:  # adopts the cached choice
        sols, tot, st = ctx.search_batch([(875859, 875859, 640, 0)])
        assert key(sols) == want


@pytest.mark.parametrize("depth", ["1", "2", "4"])
def test_walk_depth_keeps_results(depth, monkeypatch):
    """Every walk compaction depth tests the same terms: the batched engine
    (two rounds, mixed SHIFTs, resumed shifts) and the plain path give the
    reports, hash totals and SPRP test counts of depth 1."""
    ranges = RANGES + [(48620635, 48620654, 640, 0)]
    monkeypatch.setenv("AP26_WALK_DEPTH", "1")
    with gpu.Context(0) as ctx:
        base = ctx.search_batch(ranges)
        base_plain = ctx.search_k(366384, 0)
    monkeypatch.setenv("AP26_WALK_DEPTH", depth)
    with gpu.Context(0) as ctx:
        sols, tot, st = ctx.search_batch(ranges)
        psols, pst = ctx.search_k(366384, 0)
    assert key(sols) == key(base[0])
    assert [int(x) for x in tot] == [int(x) for x in base[1]]
    assert st.prp_tests == base[2].prp_tests and st.survivors == base[2].survivors
    assert key(psols) == key(base_plain[0]) and pst.prp_tests == base_plain[1].prp_tests


def test_batch_equals_plain_on_mixed_workload_sample(gpu_ctx):
    """At the benchmark's scale: the first K of the 256-unit mixed workload
    (SHIFT 0/640/1280, K from 1e6 to 1e8; 32 searched K, two rounds) give the
    same reports and per-unit hash totals on the batched engine as on the
    plain path, and the batched survivor count equals the plain one."""
    rng = np.random.default_rng(26)
    n = 256
    lens = np.clip(np.floor(np.exp(rng.uniform(0, np.log(4096), n))).astype(int), 1, 4096)
    starts = rng.integers(1_000_000, 100_000_001, n)
    shifts = rng.choice([0, 640, 1280], n)
    units, got = [], 0
    for s, L, sh in zip(starts, lens, shifts):
        ks = [K for K in range(int(s), int(s + L)) if will_search(K)][:4]
        if not ks:
            continue
        ks = ks[:32 - got]
        units.append((ks[0], ks[-1], int(sh), 0))
        got += len([K for K in range(ks[0], ks[-1] + 1) if will_search(K)])
        if got >= 32:
            break
    sols, tot, st = gpu_ctx.search_batch(units)
    plain_surv = 0
    for i, (a, b, sh, _) in enumerate(units):
        want = []
        for K in range(a, b + 1):
            if will_search(K):
                s, pst = gpu_ctx.search_k(K, sh)
                plain_surv += pst.survivors
                want += [(int(x["k"]), int(x["K"]), int(x["first"])) for x in s]
        assert key([x for x in sols if x["range_id"] == i]) == sorted(want), units[i]
        assert tot[i] == sum(f % 1000 + k for k, _, f in want), units[i]
    assert st.survivors == plain_surv and st.k_searched == got


def _expected(golden_index, K, shift):
    g = [x for x in golden_index if x["kmin"] == x["kmax"] == K and x["shift"] == shift]
    if not g:
        pytest.skip(f"reference fixture for K={K} SHIFT {shift} not written (tests/golden/make_fixtures.py --run)")
    return read_sol(os.path.join(GOLDEN, "ref_min10", g[0]["name"] + ".sol"))


def _unit_sol(tmp, i, unit, sols):
    d = tmp / str(i)
    d.mkdir()
    h = H.Host(unit[0], unit[1], unit[2], str(d), min_report=10)
    for s in sorted(sols.tolist(), key=lambda s: (s[2], s[0])):
        assert h.report(s[0], s[1], s[2], s[5], s[6]) == 0
    assert h.write_hash() == 0
    h.close()
    return read_sol(str(d / "SOL-AP26.txt"))


def test_bench_units_against_reference(golden_index, gpu_ctx, tmp_path):
    """Every K the benchmark times -- the batched leg's sample of the mixed
    workload (48 searched K over units of all three SHIFTs, across rounds)
    and the plain leg's first 24 searched K at SHIFT 640 -- plus every 256th
    searched K of the whole 4096-K plain-path unit, searched as units of one
    batched call: each unit's SOL records and hash line equal the reference
    build's (a multi-K unit's hash composed from its K's totals)."""
    units, _, ks = W.c4_sample(48)
    units = list(units) + [(K, K, 640, 0) for K in W.c3_ks(24)] + [(K, K, 640, 0) for K in W.c3_stratified(256)]
    ks = list(ks) + [[K] for K in W.c3_ks(24)] + [[K] for K in W.c3_stratified(256)]
    for u, kk in zip(units, ks):  # all fixtures present before the GPU work
        for K in kk:
            _expected(golden_index, K, u[2])
    sols, tot, st = gpu_ctx.search_batch(units)
    assert st.k_searched == sum(len(k) for k in ks)
    for i, (u, kk) in enumerate(zip(units, ks)):
        recs, T = [], 0
        for K in kk:
            r, hl = _expected(golden_index, K, u[2])
            recs += r
            T += int(hl[8:], 16)
        want = (sorted(recs), H.hash_line(u[0], u[1], H.cap_total(T)))
        assert _unit_sol(tmp_path, i, u, sols[sols["range_id"] == i]) == want, u


def test_stratified_plain_path_against_reference(golden_index, gpu_ctx, tmp_path):
    """Every 256th searched K of the 4096-K plain-path unit (SHIFT 640) on
    the plain per-K path: SOL records and hash line equal the reference
    build's single-K runs."""
    for K in W.c3_stratified(256):
        _expected(golden_index, K, 640)
    for i, K in enumerate(W.c3_stratified(256)):
        sols, _ = gpu_ctx.search_k(K, 640)
        assert _unit_sol(tmp_path, i, (K, K, 640), sols) == _expected(golden_index, K, 640), K
