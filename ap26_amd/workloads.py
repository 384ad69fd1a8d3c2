"""The synthetic work units of the benchmark and the parity fixtures.

Every list here is a pure function of a numpy seed, so the benchmark, the
fixture writer (tests/golden/make_fixtures.py) and the GPU tests name the same
K.  A work unit is the reference's `AP26 KMIN KMAX SHIFT` range
(src/common/AP26.cpp:943-992); a searched K is one that will_search admits
(src/common/AP26.cpp:649-653).
"""
import numpy as np

# 16 consecutive K from default_rng(26).integers(100_000, 1_000_001), SHIFT 640
C1_KMIN, C1_KMAX, C1_SHIFT = 875858, 875873, 640
# 4096 consecutive K from default_rng(26).integers(40_000_000, 50_000_001), SHIFT 640
C3_K0, C3_LEN, C3_SHIFT = 48620635, 4096, 640


def will_search(K):
    """K is searched unless one of 29..59 divides it (AP26.cpp:649-653)."""
    return all(K % p for p in (29, 31, 37, 41, 43, 47, 53, 59))


def searched(kmin, kmax):
    return [K for K in range(kmin, kmax + 1) if will_search(K)]


def c3_ks(n, skip=0):
    """The first skip+n searched K of the plain-path unit, minus the first skip."""
    out, K = [], C3_K0
    while len(out) < skip + n:
        if will_search(K):
            out.append(K)
        K += 1
    return out[skip:]


def c3_stratified(every=256):
    """Every `every`-th searched K of the whole 4096-K plain-path unit."""
    return searched(C3_K0, C3_K0 + C3_LEN - 1)[::every]


def c4_units(seed=26, n=256):
    """The batched workload: n units with length floor(exp(U(0, ln 4096)))
    clamped to [1, 4096], start uniform in [1e6, 1e8], SHIFT from {0, 640, 1280}."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.floor(np.exp(rng.uniform(0, np.log(4096), n))).astype(int), 1, 4096)
    starts = rng.integers(1_000_000, 100_000_001, n)
    shifts = rng.choice([0, 640, 1280], n)
    return [(int(s), int(s + L - 1), int(sh)) for s, L, sh in zip(starts, lens, shifts)]


def bitrev_order(n):
    """0..n-1 (n a power of two) in bit-reversed order: every prefix of the
    order is spread evenly over the whole list."""
    b = n.bit_length() - 1
    return [int(format(i, f"0{b}b")[::-1], 2) for i in range(n)]


def c4_visit_order(seed=26):
    """Unit indices of the batched workload in the order the sample visits
    them: bit-reversed index order (every prefix is spread over the whole
    list), taken round-robin over the SHIFT classes 0, 640, 1280 (every prefix
    of three or more units holds all three SHIFTs)."""
    units = c4_units(seed)
    order = bitrev_order(len(units))
    cls = [[i for i in order if units[i][2] == s] for s in (0, 640, 1280)]
    seq, ptr = [], [0, 0, 0]
    while len(seq) < len(units):
        for c in range(3):
            if ptr[c] < len(cls[c]):
                seq.append(cls[c][ptr[c]])
                ptr[c] += 1
    return seq


def c4_sample(steps, per_unit=3, seed=26):
    """A bounded sample of the batched workload holding `steps` searched K.

    Units are visited in c4_visit_order; each contributes its first
    `per_unit` searched K (plus what earlier short units could not give) as a
    sub-range (kmin, kmax, SHIFT, 0) of its own range, so units of several K
    run across the engine's round boundaries.  The sample for fewer steps is a
    prefix of the sample for more (the last unit possibly cut shorter), so one
    set of fixtures covers every step count up to the largest.  Returns the
    units, their unit indices and the searched K of each."""
    units = c4_units(seed)
    out, idx, ks_all, got, owed = [], [], [], 0, 0
    for i in c4_visit_order(seed):
        if got >= steps:
            break
        kmin, kmax, sh = units[i]
        quota = min(per_unit + owed, steps - got)
        ks = []
        K = kmin
        while K <= kmax and len(ks) < quota:
            if will_search(K):
                ks.append(K)
            K += 1
        owed = quota - len(ks)
        if not ks:
            continue
        out.append((ks[0], ks[-1], sh, 0))
        idx.append(i)
        ks_all.append(ks)
        got += len(ks)
    return out, idx, ks_all


def c5_units(n, seed=27, max_len=8192):
    """Serving work units: length floor(exp(U(0, ln max_len))), start uniform
    in [1e6, 1e8], SHIFT from {0, 640, 1280}."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.floor(np.exp(rng.uniform(0, np.log(max_len), n))).astype(int), 1, max_len)
    starts = rng.integers(1_000_000, 100_000_001, n)
    shifts = rng.choice([0, 640, 1280], n)
    return [(int(s), int(s + L - 1), int(sh)) for s, L, sh in zip(starts, lens, shifts)]
