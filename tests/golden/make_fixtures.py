"""Write the golden fixtures under tests/golden/ from the reference CPU build.

The reference application is compiled from its own sources by
oracle/Makefile_ref into oracle/_ref/ (AP26 with the default SOL threshold 20,
AP26_min10 with -DMINIMUM_AP_LENGTH_TO_REPORT=10, so every AP10+ record is
written).  Each run below executes `AP26_min10 KMIN KMAX SHIFT` in its own
directory under oracle/_ref/runs/<name>/ and this script copies the resulting
SOL-AP26.txt (records + hash line), the final state file and the Checkpoint /
Non-Solution lines it printed into tests/golden/ref_min10/.

It also copies the reference's own golden files tests/TEST-*.txt (data).

    python tests/golden/make_fixtures.py [--run]   # --run: execute missing runs
"""
import json
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from ap26_amd import workloads as W  # noqa: E402
REF = "/root/reference"
REF_OUT = os.path.join(ROOT, "oracle", "_ref")
RUNS = os.path.join(REF_OUT, "runs")
DST = os.path.join(HERE, "ref_min10")

# name, KMIN, KMAX, SHIFT.  c1_*: the 13 searched K of 875858..875873 (16
# consecutive K from numpy default_rng(26).integers(100000, 1000001)) at
# SHIFT 640; c3_*: K from the 4096-K range starting at
# default_rng(26).integers(40000000, 50000001) = 48620635; s1280_*: SHIFT 1280
# at both ends of the batched K range; p61_*: a K divisible by 61, where
# STEP = 0 (mod 61); h4_*: K = 61*67*71*73, the most sieve primes a searched
# K <= 6e8 can hold, so the sieve keeps about 5.4x the usual survivors;
# range_*: a multi-K unit with a skipped K; k6e8_*: the largest K the build
# admits (6e8, where STEP = K*23# is largest) at SHIFT 640 and 1280; mid_*: a
# K between the config ranges and 6e8; s5120_*: a SHIFT beyond the configs'.
RUN_LIST = [
    ("t366384", 366384, 366384, 0),
    ("t44121555", 44121555, 44121555, 0),
    ("t47715109", 47715109, 47715109, 0),
] + [(f"c1_{K}", K, K, 640) for K in (875859, 875860, 875861, 875862, 875863, 875865, 875866,
                                       875868, 875869, 875870, 875871, 875872, 875873)] + [
    ("c3_48620635", 48620635, 48620635, 640),
    ("c3_48622684", 48622684, 48622684, 640),
    ("s1280_1000003", 1000003, 1000003, 1280),
    ("s1280_99999989", 99999989, 99999989, 1280),
    ("p61_976061", 976061, 976061, 0),
    ("range_875858_875860", 875858, 875860, 640),
    ("h4_21182921", 21182921, 21182921, 0),
    ("k6e8_600000000", 600000000, 600000000, 640),
    ("k6e8_599999999", 599999999, 599999999, 1280),
    ("mid_300000007", 300000007, 300000007, 0),
    ("s5120_1000003", 1000003, 1000003, 5120),
]
# bench_*: every searched K the benchmark times -- the batched leg's sample of
# the 256-unit workload for up to 48 steps (ap26_amd.workloads.c4_sample, each
# K at its unit's SHIFT) and the plain leg's first 24 searched K of the
# 4096-K unit at SHIFT 640; c3s_*: every 256th searched K of that whole unit.
_have = {r[1] for r in RUN_LIST if r[1] == r[2]}
for _u, _ks in zip(*W.c4_sample(48)[::2]):
    RUN_LIST += [(f"bench_{K}_{_u[2]}", K, K, _u[2]) for K in _ks]
RUN_LIST += [(f"bench_{K}_640", K, K, 640) for K in W.c3_ks(24) if K not in _have]
RUN_LIST += [(f"c3s_{K}", K, K, 640) for K in W.c3_stratified(256) if K not in _have]


def run_one(name, kmin, kmax, shift):
    d = os.path.join(RUNS, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "stdout.txt"), "w") as out, \
            open(os.path.join(d, "stderr.txt"), "w") as err:
        subprocess.run([os.path.join(REF_OUT, "AP26_min10"), str(kmin), str(kmax), str(shift)],
                       cwd=d, stdout=out, stderr=err, check=True)


def collect(name, kmin, kmax, shift):
    d = os.path.join(RUNS, name)
    sol = os.path.join(d, "SOL-AP26.txt")
    if not os.path.exists(sol):
        return None
    lines = open(sol).read().splitlines()
    if not lines or len(lines[-1]) != 16:
        return None  # unfinished run
    shutil.copy(sol, os.path.join(DST, f"{name}.sol"))
    stdout = open(os.path.join(d, "stdout.txt")).read().splitlines()
    ckpt = [l for l in stdout if l.startswith("Checkpoint:")]
    nonsol = [l for l in stdout if l.startswith("Non-Solution:")]
    secs = [int(l.split()[-2]) for l in stdout if l.startswith("Computation of K:")]
    states = {}
    for s in ("AP26-state.a.txt", "AP26-state.b.txt"):
        p = os.path.join(d, s)
        if os.path.exists(p):
            states[s] = open(p).read()
    return {"name": name, "kmin": kmin, "kmax": kmax, "shift": shift,
            "records": len(lines) - 1, "hash": lines[-1], "checkpoints": ckpt,
            "nonsolutions": nonsol, "seconds_per_k": secs, "state_files": states,
            "min_report": 10}


def main():
    os.makedirs(DST, exist_ok=True)
    tdst = os.path.join(HERE, "ref_tests")
    os.makedirs(tdst, exist_ok=True)
    for f in sorted(os.listdir(os.path.join(REF, "tests"))):
        if f.startswith("TEST-") and f.endswith(".txt"):
            shutil.copy(os.path.join(REF, "tests", f), os.path.join(tdst, f))
    old = {}
    if os.path.exists(os.path.join(DST, "index.json")):
        with open(os.path.join(DST, "index.json")) as f:
            old = {m["name"]: m for m in json.load(f)}
    if "--run" in sys.argv:
        # runs whose fixture is already committed are not repeated
        todo = [r for r in RUN_LIST if r[0] not in old and
                not os.path.exists(os.path.join(RUNS, r[0], "SOL-AP26.txt"))]
        jobs = int(os.environ.get("AP26_FIXTURE_JOBS", "6"))
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            list(ex.map(lambda r: run_one(*r), todo))
    index = []
    for r in RUN_LIST:
        meta = collect(*r) or old.get(r[0])
        if meta:
            index.append(meta)
        else:
            print("missing run", r[0])
    with open(os.path.join(DST, "index.json"), "w") as f:
        json.dump(index, f, indent=1)
    print(f"{len(index)} runs collected into {DST}")


if __name__ == "__main__":
    main()
