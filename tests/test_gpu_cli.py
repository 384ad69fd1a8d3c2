"""The AP26 command (ap26_amd/bin/AP26) as a drop-in for the reference's
`AP26 KMIN KMAX SHIFT`: SOL file, hash line, state files, resume."""
import os
import subprocess

import pytest

from ap26_amd import host as H
from conftest import GOLDEN, ROOT, read_sol

pytestmark = pytest.mark.gpu
AP26 = os.environ.get("AP26_BIN", os.path.join(ROOT, "ap26_amd", "bin", "AP26"))


def run(args, cwd, env=None):
    full = None if env is None else {**os.environ, **env}
    r = subprocess.run([AP26, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=300, env=full)
    return r


def golden(golden_index, name):
    return [g for g in golden_index if g["name"] == name][0]


def test_test_file_byte_identical(tmp_path):
    """README.md:47-51: SOL-AP26.txt of `AP26 366384 366384 0` equals tests/TEST-366384.txt."""
    r = run([366384, 366384, 0], tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "SOL-AP26.txt").read_text() == \
        open(os.path.join(GOLDEN, "ref_tests", "TEST-366384.txt")).read()
    states = [(tmp_path / f).read_text().split() for f in ("AP26-state.a.txt", "AP26-state.b.txt")]
    final = [s for s in states if s[3] == "366385"][0]
    assert final[:6] == ["366384", "366384", "0", "366385", "0", str(0x15550)]


def _units_hash(golden_index, names, kmin, kmax):
    T = 0
    recs = []
    for n in names:
        rr, hl = read_sol(os.path.join(GOLDEN, "ref_min10", n + ".sol"))
        recs += rr
        T += int(hl[8:], 16)
    return sorted(recs), H.hash_line(kmin, kmax, H.cap_total(T))


def test_multi_k_unit_with_skipped_k(golden_index, tmp_path):
    r = run([875858, 875860, 640, "-report", 10], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "2 K in this range remaining to be searched (1 skipped, 0 done)" in r.stdout
    want = _units_hash(golden_index, ["c1_875859", "c1_875860"], 875858, 875860)
    assert read_sol(str(tmp_path / "SOL-AP26.txt")) == want


def test_resume_from_reference_style_checkpoint(golden_index, tmp_path):
    """A state file as the reference writes it after finishing K=875859:
    the resumed run searches only 875860 and ends with the full unit's hash."""
    T59 = int(golden(golden_index, "c1_875859")["hash"][8:], 16)
    (tmp_path / "AP26-state.a.txt").write_text(f"875858 875860 640 875860 0 {T59} 1234.000000\n")
    r = run([875858, 875860, 640, "-report", 10], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Resuming search from checkpoint." in r.stdout
    recs, hline = read_sol(str(tmp_path / "SOL-AP26.txt"))
    want_recs, want_hash = _units_hash(golden_index, ["c1_875859", "c1_875860"], 875858, 875860)
    assert hline == want_hash
    assert recs == read_sol(os.path.join(GOLDEN, "ref_min10", "c1_875860.sol"))[0]


def test_resume_mid_k_and_shift_checkpoints(golden_index, tmp_path):
    """-shiftckpt checkpoints at ITER 4 and 8 inside each K: the Checkpoint
    lines equal those the reference AVX2 build printed for the same unit; a
    run resumed from an ITER=4 state ends with the golden hash."""
    a = tmp_path / "a"
    a.mkdir()
    r = run([875861, 875861, 640, "-report", 10, "-shiftckpt"], a)
    assert r.returncode == 0, r.stderr
    ck = [l for l in r.stdout.splitlines() if l.startswith("Checkpoint:")]
    g = golden(golden_index, "c1_875861")
    assert ck == g["checkpoints"]
    assert [l.split("ITER:")[1] for l in ck] == ["0", "4", "8", "0"]
    assert read_sol(str(a / "SOL-AP26.txt")) == read_sol(os.path.join(GOLDEN, "ref_min10", "c1_875861.sol"))
    # replay: take the hash the run had after 4 shifts, resume from there
    b = tmp_path / "b"
    b.mkdir()
    # the hash after shifts 0..3, from a 4-shift run of the same unit
    c = tmp_path / "c"
    c.mkdir()
    from ap26_amd import gpu
    with gpu.Context(0) as ctx:
        sols, _ = ctx.search_span(875861, 640, 0, 4, 0, 10840)
    hh = H.Host(875861, 875861, 640, str(c), min_report=10)
    for s in sols.tolist():
        assert hh.report(s[0], s[1], s[2]) == 0
    h4 = hh.hash
    hh.close()
    (b / "AP26-state.b.txt").write_text(f"875861 875861 640 875861 4 {h4} 1234.000000\n")
    r = run([875861, 875861, 640, "-report", 10], b)
    assert r.returncode == 0, r.stderr
    assert read_sol(str(b / "SOL-AP26.txt"))[1] == g["hash"]


def test_bad_arguments(tmp_path):
    r = run([1], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stdout


def test_batch_flag_same_output(golden_index, tmp_path):
    r = run([875858, 875860, 640, "-report", 10, "-batch"], tmp_path)
    assert r.returncode == 0, r.stderr
    want = _units_hash(golden_index, ["c1_875859", "c1_875860"], 875858, 875860)
    assert read_sol(str(tmp_path / "SOL-AP26.txt")) == want


def test_batch_flag_chunks_equal_per_k_run(tmp_path):
    """`-batch` runs the unit in chunks of searched K (here one K per call) and
    replays each K's checkpoint and reports: SOL file and state files are byte
    for byte those of the per-K run, also when resuming mid-unit."""
    unit = [875858, 875862, 640, "-report", 10]
    outs = {}
    for name, extra, env in (("perk", [], None), ("batch1", ["-batch"], {"AP26_BATCH_K_PER_CALL": "1"}),
                             ("batch2", ["-batch"], {"AP26_BATCH_K_PER_CALL": "2"})):
        d = tmp_path / name
        d.mkdir()
        r = run(unit + extra, d, env)
        assert r.returncode == 0, (name, r.stderr)
        outs[name] = ((d / "SOL-AP26.txt").read_bytes(), _state_fields(d))
    assert outs["batch1"] == outs["perk"]
    assert outs["batch2"] == outs["perk"]
    # resume at K=875861 from a checkpoint written by the per-K run's sequence
    sol = outs["perk"][0].decode().splitlines()
    T = 0
    for line in sol[:-1]:
        k, K, first = (int(x) for x in line.split())
        if K < 875861:
            T += (first % 1000) + k
    res = {}
    for name, extra, env in (("perk", [], None), ("batch", ["-batch"], {"AP26_BATCH_K_PER_CALL": "1"})):
        d = tmp_path / ("resume_" + name)
        d.mkdir()
        (d / "AP26-state.a.txt").write_text(f"875858 875862 640 875861 0 {T} 1234.000000\n")
        r = run(unit + extra, d, env)
        assert r.returncode == 0, (name, r.stderr)
        res[name] = ((d / "SOL-AP26.txt").read_bytes(), _state_fields(d))
    assert res["batch"] == res["perk"]


def _state_fields(d):
    """State files without the trickle timestamp (the reference's last field)."""
    out = {}
    for f in ("AP26-state.a.txt", "AP26-state.b.txt"):
        p = d / f
        out[f] = p.read_text().split()[:6] if p.exists() else None
    return out


def test_manifest_units_match_single_runs(golden_index, tmp_path):
    """`AP26 -manifest FILE` runs several units on the batched engine and
    leaves in unit_<KMIN>_<KMAX>_<SHIFT>/ the same SOL file (records and hash
    line, byte for byte) and state files as `AP26 KMIN KMAX SHIFT` run alone in
    that directory; one unit resumes from a checkpoint left by an earlier run."""
    units = [(366384, 366384, 0), (875858, 875860, 640), (976061, 976062, 0), (29, 31, 0)]
    single = tmp_path / "single"
    single.mkdir()
    for a, b, sh in units:
        d = single / f"unit_{a}_{b}_{sh}"
        d.mkdir()
        if (a, b) == (875858, 875860):  # resumed after K=875859, as the reference writes it
            T59 = int(golden(golden_index, "c1_875859")["hash"][8:], 16)
            (d / "AP26-state.a.txt").write_text(f"875858 875860 640 875860 0 {T59} 1234.000000\n")
        r = run([a, b, sh, "-report", 10], d)
        assert r.returncode == 0, r.stderr
    batch = tmp_path / "batch"
    batch.mkdir()
    d0 = batch / "unit_875858_875860_640"
    d0.mkdir()
    T59 = int(golden(golden_index, "c1_875859")["hash"][8:], 16)
    (d0 / "AP26-state.a.txt").write_text(f"875858 875860 640 875860 0 {T59} 1234.000000\n")
    (batch / "units.txt").write_text("# KMIN KMAX SHIFT\n" + "".join(f"{a} {b} {sh}\n" for a, b, sh in units))
    r = run(["-manifest", "units.txt", "-report", 10], batch)
    assert r.returncode == 0, r.stderr
    for a, b, sh in units:
        name = f"unit_{a}_{b}_{sh}"
        s_sol = (single / name / "SOL-AP26.txt").read_text()
        b_sol = (batch / name / "SOL-AP26.txt").read_text()
        assert sorted(b_sol.splitlines()[:-1]) == sorted(s_sol.splitlines()[:-1]), name
        assert b_sol.splitlines()[-1] == s_sol.splitlines()[-1], name
        assert _state_fields(batch / name) == _state_fields(single / name), name
    # and against the reference's own runs
    want = read_sol(os.path.join(GOLDEN, "ref_min10", "t366384.sol"))
    assert read_sol(str(batch / "unit_366384_366384_0" / "SOL-AP26.txt")) == want


def test_metrics_records(tmp_path):
    """-metrics: one progress record per checkpoint (the trickle-up fields),
    progress non-decreasing to 1.0; the non-BOINC progress line on stdout."""
    import json
    r = run([875858, 875862, 640, "-report", 10, "-metrics", "m.jsonl"], tmp_path)
    assert r.returncode == 0, r.stderr
    recs = [json.loads(l) for l in open(tmp_path / "m.jsonl")]
    ks = [K for K in range(875858, 875863) if H.will_search(K)]
    assert [x["K"] for x in recs] == ks + [875863]
    prog = [x["progress"] for x in recs]
    assert prog == sorted(prog) and prog[0] == 0.0 and prog[-1] == 1.0
    assert recs[-1]["result_hash"] == int((tmp_path / "SOL-AP26.txt").read_text().splitlines()[-1][8:], 16)
    assert "progress: 100.00%" in r.stdout
    m = tmp_path / "m2"
    m.mkdir()
    (m / "u.txt").write_text("875858 875862 640\n366384 366384 0\n")
    r = run(["-manifest", "u.txt", "-report", 10, "-metrics", "m.jsonl", "-q"], m)
    assert r.returncode == 0, r.stderr
    for d in ("unit_875858_875862_640", "unit_366384_366384_0"):
        recs = [json.loads(l) for l in open(m / d / "m.jsonl")]
        assert recs[-1]["progress"] == 1.0 and recs[-1]["k_done"] == recs[-1]["k_count"]


def test_walk_limit_exits_like_the_reference(tmp_path):
    """The reference CPU build exits (status 1, "modulus out of range") when
    a walked term reaches 2^62 (src/common/PrimeQ_x86.h:160-166).  K = 6e8
    at SHIFT 16912 is refused before the search; at SHIFT 14000 the walks
    reach 2^62 on the device; -batch stops at the same K with the same state
    files as the per-K run."""
    for shift in (16912, 14000):
        d = tmp_path / str(shift)
        d.mkdir()
        r = run([600000000, 600000000, shift], d)
        assert r.returncode == 1 and "modulus out of range" in r.stderr, r.stderr
        assert "Workunit complete." not in r.stderr
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    ra = run([599999998, 600000000, 14000], a)
    rb = run([599999998, 600000000, 14000, "-batch"], b)
    assert ra.returncode == rb.returncode == 1
    for f in ("AP26-state.a.txt", "AP26-state.b.txt"):
        fa, fb = a / f, b / f
        assert fa.exists() == fb.exists()
        if fa.exists():
            assert fa.read_text().split()[:6] == fb.read_text().split()[:6]


def test_config1_whole_unit(golden_index, tmp_path):
    """The CPU-runnable configuration as one unit: `AP26 875858 875873 640
    -report 10` (16 consecutive K, 13 searched) writes the records of the
    reference build's 13 single-K runs and the unit hash composed from their
    totals."""
    r = run([875858, 875873, 640, "-report", 10], tmp_path)
    assert r.returncode == 0, r.stderr
    names = [f"c1_{K}" for K in range(875858, 875874) if H.will_search(K)]
    assert len(names) == 13
    assert read_sol(str(tmp_path / "SOL-AP26.txt")) == _units_hash(golden_index, names, 875858, 875873)
