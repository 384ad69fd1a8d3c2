"""Write tests/golden/n43_report_index.json: for two K of the golden set, the
n43 starts (index j of the n43 enumeration, src/opencl/AP26.h:125-136) whose
search produces each report, found with the GPU library by bisection over
spans of n43 starts.  The oracle test that reads it searches exactly those
starts on the CPU and compares the reports with the reference build's
records, which pins the oracle's n43 values and order without a full-K
oracle run.  Run on a machine with the GPU:

    python tests/golden/make_n43_index.py [--out PATH]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ap26_amd import gpu  # noqa: E402

UNITS = [(366384, 0, "t366384"), (875859, 640, "c1_875859")]
CHUNK = 64


def reports(ctx, K, sh, j0, j1):
    sols, _ = ctx.search_span(K, sh, 0, 10, j0, j1)
    return sorted((int(s["k"]), int(s["first"])) for s in sols)


def locate(ctx, K, sh, j0, j1, out):
    r = reports(ctx, K, sh, j0, j1)
    if not r:
        return
    if j1 - j0 == 1:
        out.append({"j": j0, "reports": [list(x) for x in r]})
        return
    mid = (j0 + j1) // 2
    locate(ctx, K, sh, j0, mid, out)
    locate(ctx, K, sh, mid, j1, out)


def main():
    res = []
    with gpu.Context(0) as ctx:
        for K, sh, name in UNITS:
            idx = []
            for c in range(0, 10840, CHUNK):
                locate(ctx, K, sh, c, min(10840, c + CHUNK), idx)
            res.append({"K": K, "shift": sh, "golden": name, "starts": idx})
            print(K, sh, len(idx), "starts with reports", flush=True)
    out = os.path.join(HERE, "n43_report_index.json")
    if "--out" in sys.argv:
        out = sys.argv[sys.argv.index("--out") + 1]
    with open(out, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
