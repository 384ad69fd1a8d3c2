"""Golden vectors for the reference's ReportSolution (src/common/AP26.cpp:549-602),
including the Non-Solution recursion that no shipped fixture reaches.

Runs oracle/_ref/ref_report_harness (the reference host driver compiled from
its sources, see oracle/Makefile_ref) on synthetic reports and writes
tests/golden/report_vectors.json: for each case the inputs, the result_hash
after the call (starting from 0), the printed Non-Solution lines, the SOL
lines written (default threshold 20) and the exit status.

    python tests/golden/make_report_vectors.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_report_harness")

# (k, K, first): 23#-APs built around base-2 strong pseudoprimes
CASES = [
    # prime, spsp(2) failing base 13, prime
    [(3, 23, 2147171762737)],
    # a lone spsp(2) failing base 11
    [(1, 1, 3215031751)],
    # spsp(2) failing base 17 between two primes
    [(3, 35, 3466941409933)],
    # 3825123056546413051 passes all nine bases: reported as a solution
    [(3, 14, 3825123053423112871)],
    # a real AP25 from tests/TEST-366384.txt: SOL line written
    [(25, 366384, 6171054912832631)],
    # sub-AP of the AP25 and its hash only (k < 20)
    [(12, 366384, 6171054912832631)],
    # several reports in sequence
    [(3, 23, 2147171762737), (25, 366384, 6171054912832631), (1, 1, 3215031751)],
    # not a base-2 PRP at all: "Computation error", exit status 1
    [(2, 1, 15)],
    # even first term: validate returns 0 immediately, so it is a Non-Solution
    # whose base-2 check also fails -> exit 1
    [(2, 1, 16)],
]


def main():
    out = []
    for case in CASES:
        with tempfile.TemporaryDirectory() as d:
            args = [str(x) for rep in case for x in rep]
            r = subprocess.run([HARNESS, *args], cwd=d, capture_output=True, text=True)
            hashes = [int(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("hash ")]
            nonsol = [l for l in r.stdout.splitlines() if l.startswith("Non-Solution:")]
            solp = os.path.join(d, "SOL-AP26.txt")
            sol = open(solp).read().splitlines() if os.path.exists(solp) else []
            out.append({"reports": case, "hashes": hashes, "nonsolutions": nonsol,
                        "sol_lines": sol, "exit": r.returncode})
    with open(os.path.join(HERE, "report_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
