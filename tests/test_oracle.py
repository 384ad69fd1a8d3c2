"""Pins the oracle (oracle/ap26_oracle.c) before it is trusted as the checker:
golden files written by the reference CPU build, the reference's own
ReportSolution run on synthetic reports, brute force from the definitions,
and textbook number theory.  CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, read_sol

MOD = 258559632607830
PRIM23 = 223092870


def primes_between(a, b):
    return [p for p in range(a, b + 1) if p > 1 and all(p % d for d in range(2, int(p ** 0.5) + 1))]


SIEVE_PRIMES = primes_between(61, 541)

# OEIS A001262: strong pseudoprimes to base 2 below 2^16
SPSP2_BELOW_2_16 = [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281]


def is_prime_td(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    d = 3
    while d * d <= n:
        if n % d == 0:
            return False
        d += 2
    return True


def sprp_pow(a, n):
    """Miller-Rabin round with Python big-int pow: an independent implementation."""
    if n < 3 or n % 2 == 0:
        return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(a, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def test_prime_tables():
    assert O.primes() == SIEVE_PRIMES
    offs = O.offsets()
    assert offs[0] == 0 and offs[-1] + 541 == 23693
    assert all(offs[i + 1] - offs[i] == SIEVE_PRIMES[i] for i in range(82))


def test_will_search():
    assert O.will_search(366384) and O.will_search(44121555) and O.will_search(47715109)
    for p in (29, 31, 37, 41, 43, 47, 53, 59):
        assert not O.will_search(p * 12345)
    # the unit 875858..875873 of the golden set: 13 of 16 searched
    assert [K for K in range(875858, 875874) if O.will_search(K)] == [
        875859, 875860, 875861, 875862, 875863, 875865, 875866, 875868, 875869, 875870,
        875871, 875872, 875873]


@pytest.mark.parametrize("K", [1, 366384, 875859, 17835, 17836, 44121555, 48620635, 99999989,
                               100004095, 600000000])
def test_kconsts_match_reference_split_formula(K):
    """The reference keeps int64 by splitting K = 17835q + r (AP26.h:112-120);
    the oracle uses exact 128-bit products.  Evaluate the split form with int64
    wrap-around and compare."""
    def i64(x):
        return ((x + 2 ** 63) % 2 ** 64) - 2 ** 63

    def split(c, add=0):
        v = i64(i64(c * (K % 17835)) + i64(((c * 17835) % MOD) * (K // 17835)) + add)
        return v % MOD

    kc = O.kconsts(K)
    assert kc.STEP == K * PRIM23
    assert kc.n0 == split(106990415896110, 94805198622871)
    pres = [16681266619860, 181690552643340, 132432982555230, 126273308948010,
            115526644356690, 48784836341100, 100794433050510]
    got = [kc.S31, kc.S37, kc.S41, kc.S43, kc.S47, kc.S53, kc.S59]
    assert got == [split(c) for c in pres]


def test_n43_table_shape():
    t = O.n43_table(366384)
    assert len(t) == 10840 and len(set(t.tolist())) == 10840
    assert t.min() >= 0 and t.max() < MOD
    # the reference's 10840 = 1355 guarded (i31, i37, i41) triples x 2 x 4
    assert all(v % 2 == 1 for v in t[:50])


@pytest.mark.parametrize("K", [366384, 976061, 44121555])
def test_ok_table_definition(K):
    """OK_p[r] == 0 iff a term r + i*STEP (0 <= i <= 23) is divisible by p."""
    ok = O.ok_table(K)
    offs = O.offsets()
    step = K * PRIM23
    for i, p in enumerate(SIEVE_PRIMES):
        for r in range(p):
            bad = any((r + j * step) % p == 0 for j in range(24))
            assert ok[offs[i] + r] == (0 if bad else 1), (p, r)


@pytest.mark.parametrize("K,shift", [(366384, 0), (875859, 640 + 64 * 7), (976061, 1280)])
def test_okok_table_definition(K, shift):
    ok = O.ok_table(K)
    okok = O.okok_table(K, shift)
    offs = O.offsets()
    rng = np.random.default_rng(K)
    for i in rng.choice(83, 12, replace=False):
        p = SIEVE_PRIMES[i]
        for r in rng.choice(p, 5, replace=False):
            w = 0
            for jj in range(64):
                w |= int(ok[offs[i] + (int(r) + (jj + shift) * MOD) % p]) << jj
            assert int(okok[offs[i] + r]) == w


def test_sprp_small_exhaustive():
    """Odd n < 2^16: base-2 SPRP iff prime or one of the known strong pseudoprimes."""
    for n in range(3, 1 << 16, 2):
        expect = is_prime_td(n) or n in SPSP2_BELOW_2_16
        assert O.prp2(n) == expect, n
    assert not O.prp2(1) and not O.prp2(0) and not O.prp2(4)


def test_sprp_large_against_pow():
    rng = np.random.default_rng(7)
    xs = (rng.integers(2 ** 40, 2 ** 63, size=4000, dtype=np.uint64) | np.uint64(1)).tolist()
    xs += [2 ** 61 - 1, 2 ** 62 - 57, 2 ** 63 - 25, 2 ** 64 - 59, 2 ** 64 - 1, 3825123056546413051,
           341550071728321, 3474749660383, 2152302898747, 3215031751]
    for n in xs:
        for a in (2, 3, 23):
            assert O.sprp(a, n) == sprp_pow(a, n), (a, n)


def test_walk_semantics_on_reference_ap25():
    """The AP25 of tests/TEST-366384.txt: the walk from n = first + (j-5)*STEP
    reports (25, first) iff the forward run (25 - j terms) reaches 10."""
    K, first = 366384, 6171054912832631
    step = K * PRIM23
    for j in range(25):
        k, f = O.walk(first + (j - 5) * step, step)
        if 25 - j >= 10:
            assert (k, f) == (25, first), j
        else:
            assert k == 25 - j and k < 10


def test_report_vectors_from_reference():
    """ReportSolution of the reference (run through oracle/_ref/ref_report_harness)
    on synthetic reports, Non-Solution recursion included."""
    vecs = json.load(open(os.path.join(GOLDEN, "report_vectors.json")))
    for v in vecs:
        h = 0
        recs_all = []
        rc = 0
        for (k, K, first), want in zip(v["reports"], v["hashes"] + [None] * len(v["reports"])):
            rc, h, recs, nonsol = O.report(k, K, first, h, min_report=20)
            if rc:
                break
            recs_all += [f"{r['k']} {r['K']} {r['first']}" for r in recs]
            assert h == want
        assert (rc != 0) == (v["exit"] != 0)
        if rc == 0:
            assert recs_all == v["sol_lines"]


def test_hash_word_matches_golden_files(golden_index):
    """Low 32 bits = sum of (first % 1000 + k) over all AP10+ reports; high 32
    bits = KMIN + KMAX (AP26.cpp:295-357)."""
    for g in golden_index:
        recs, hline = read_sol(os.path.join(GOLDEN, "ref_min10", g["name"] + ".sol"))
        assert not g["nonsolutions"]
        total = sum(f % 1000 + k for k, _, f in recs)
        assert "%016X" % O.hash_word(g["kmin"], g["kmax"], total) == hline
    for name in ("366384", "44121555", "47715109"):
        lines = open(os.path.join(GOLDEN, "ref_tests", f"TEST-{name}.txt")).read().splitlines()
        g = [x for x in golden_index if x["name"] == "t" + name][0]
        assert lines[-1] == g["hash"]
        min10, _ = read_sol(os.path.join(GOLDEN, "ref_min10", f"t{name}.sol"))
        assert sorted(tuple(int(x) for x in l.split()) for l in lines[:-1]) == \
            [r for r in min10 if r[0] >= 20]


def _candidates(K, shift0, nshift, j, nblocks):
    """Brute-force candidate set of the first nblocks n53 blocks of n43 index j."""
    kc = O.kconsts(K)
    n43 = int(O.n43_table(K)[j])
    pos = []
    for blk in range(nblocks):
        i43, rem = divmod(blk, 667)
        i47, i53 = divmod(rem, 29)
        n53 = (n43 + i43 * kc.S43 + i47 * kc.S47 + i53 * kc.S53) % MOD
        for i59 in range(35):
            pos.append((n53 + i59 * kc.S59) % MOD)
    pos = np.array(pos, dtype=np.int64)
    cands = []
    for s in range(nshift):
        for jj in range(64):
            cands.append(pos + (shift0 + 64 * s + jj) * MOD)
    return pos, np.concatenate(cands), kc.STEP


@pytest.mark.parametrize("K,shift", [(366384, 0), (976061, 640)])
def test_sieve_brute_force(K, shift):
    """Survivors of the first 58 n53 blocks of one n43 start equal the set
    defined from scratch: no factor 61..541 in any of the 24 terms n + i*STEP
    and no factor 7..23 in n."""
    pos, c, step = _candidates(K, shift, 2, 5, 58)
    keep = np.ones(len(c), dtype=bool)
    for q in (7, 11, 13, 17, 19, 23):
        keep &= (c % q) != 0
    for p in SIEVE_PRIMES:
        cp = c % p
        sp = step % p
        for i in range(24):
            keep &= (cp + i * sp) % p != 0
    want = np.sort(c[keep])
    got = O.sieve_span(K, shift, 0, 2, 5, 6)
    got = np.sort(got[np.isin(got % MOD, pos)])
    assert np.array_equal(got, want)


def test_sieve_span_shift_union():
    """Shifts are independent: a 10-shift span is the union of 1-shift spans."""
    whole = np.sort(O.sieve_span(875859, 640, 0, 10, 300, 301))
    parts = np.sort(np.concatenate([O.sieve_span(875859, 640, s, s + 1, 300, 301) for s in range(10)]))
    assert np.array_equal(whole, parts)


@pytest.mark.slow
def test_oracle_full_k_against_reference_golden():
    """Full K=366384, SHIFT 0 with 8 threads (~3 min): the reference's own
    survivor count (src/opencl/AP26.h:343) and its min10 SOL file + hash."""
    sols, nsurv = O.search_span(366384, 0, 0, 10, 0, 10840, threads=8)
    assert nsurv == 38838420
    h = 0
    recs = []
    for s in sols:
        rc, h, r, _ = O.report(int(s["k"]), int(s["K"]), int(s["first"]), h, min_report=10)
        assert rc == 0
        recs += [(int(x["k"]), int(x["K"]), int(x["first"])) for x in r]
    want, hline = read_sol(os.path.join(GOLDEN, "ref_min10", "t366384.sol"))
    assert sorted(recs) == want
    assert "%016X" % O.hash_word(366384, 366384, h) == hline


# The reference GPU kernel's early-exit grouping of the 83 sieve primes
# (src/opencl/kernels/sieve.cl:32-114): a mask is ANDed over a group and
# tested once per group; groups of 5, 5, 5, 4, 4 primes, then pairs.
REFERENCE_GROUPS = [5, 5, 5, 4, 4] + [2] * 30


@pytest.mark.parametrize("K,shift,it,j", [(875859, 640, 3, 123), (366384, 0, 9, 10000)])
def test_lookup_count_against_independent_count(K, shift, it, j):
    """The sieve's work count behind its roofline (ap26o_lookup_count)
    equals a count written here with numpy: every position of one n43 start
    (n59 = n43 + i43*S43 + i47*S47 + i53*S53 + i59*S59 mod MOD, 443,555 of
    them) is ANDed group by group over the oracle's OKOK table (pinned by
    test_okok_table_definition) and charged each group it enters."""
    assert sum(REFERENCE_GROUPS) == len(SIEVE_PRIMES) == 83
    kc = O.kconsts(K)
    n43 = int(O.n43_table(K)[j])
    i43, i47, i53, i59 = np.meshgrid(np.arange(19), np.arange(23), np.arange(29), np.arange(35), indexing="ij")
    n59 = (n43 + i43 * kc.S43 + i47 * kc.S47 + i53 * kc.S53 + i59 * kc.S59).ravel() % MOD
    tab = O.okok_table(K, shift + 64 * it)
    off = np.cumsum([0] + SIEVE_PRIMES[:-1])
    m = np.full(n59.size, np.uint64(2 ** 64 - 1), dtype=np.uint64)
    alive = np.ones(n59.size, dtype=bool)
    count, i = 0, 0
    for g in REFERENCE_GROUPS:
        count += g * int(alive.sum())
        for p, o in zip(SIEVE_PRIMES[i:i + g], off[i:i + g]):
            m &= tab[o + n59 % p]
        alive &= m != 0
        i += g
    assert count == O.lookup_count(K, shift, it, it + 1, j, j + 1)
    assert 18 * n59.size < count < 21 * n59.size  # ~18-19.5 lookups per mask


def test_oracle_n43_starts_reproduce_reference_records(golden_index):
    """The oracle's n43 values and order, without a full-K run: for two K of
    the golden set, the n43 starts that carry reports (index j of the
    enumeration, tests/golden/n43_report_index.json) are searched on the CPU
    over all 10 shifts; at the first, a middle and the last of them, and at
    a start of no report, the oracle's reports are exactly the index's, and
    every one is a record of the reference build's SOL file (k >= 10)."""
    with open(os.path.join(GOLDEN, "n43_report_index.json")) as f:
        index = json.load(f)
    assert len(index) == 2
    for u in index:
        recs, _ = read_sol(os.path.join(GOLDEN, "ref_min10", u["golden"] + ".sol"))
        rec_set = {(k, first) for k, _, first in recs}
        starts = u["starts"]
        idx_pairs = {(k, f) for s in starts for k, f in s["reports"]}
        assert {p for p in idx_pairs if p[0] >= 10} == rec_set  # the index covers every record
        with_reports = {s["j"] for s in starts}
        empty = next(j for j in range(10840) if j not in with_reports)
        for s in (starts[0], starts[len(starts) // 2], starts[-1], {"j": empty, "reports": []}):
            sols, _ = O.search_span(u["K"], u["shift"], 0, 10, s["j"], s["j"] + 1, threads=4)
            got = sorted((int(x["k"]), int(x["first"])) for x in sols)
            assert got == sorted(tuple(r) for r in s["reports"]), (u["K"], s["j"])
            assert all((k, f) in rec_set for k, f in got if k >= 10)
