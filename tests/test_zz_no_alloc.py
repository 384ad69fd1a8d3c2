"""ap26_search_span_noalloc, the entry point for callers that cannot allocate:
through the C ABI, its reports and counts equal ap26_search_span's on calls
in a row (the scratch array keeps the previous call's contents), the second
one smaller; spans beyond its 64 n43 starts are refused."""
import ctypes

import numpy as np
import pytest

from ap26_amd import gpu

pytestmark = pytest.mark.gpu


def _call(lib, fn, ctx, args, cap=4096):
    out = np.zeros(cap, dtype=gpu.SOL_DTYPE)
    n = ctypes.c_int32()
    st = gpu.Stats()
    rc = getattr(lib, fn)(ctx.h, *args, out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n),
                          ctypes.byref(st))
    return rc, sorted((int(x["k"]), int(x["K"]), int(x["first"]), int(x["shift_iter"]), int(x["prp2_mask"]),
                       int(x["valid_mask"])) for x in out[:n.value]), st


# the largest span first, then smaller ones; a K divisible by 61, 67, 71 and
# 73 (5.45x the usual survivors) at the full 64 n43 starts
SPANS = [(366384, 0, 0, 10, 0, 64), (875859, 640, 2, 7, 100, 120), (21182921, 0, 0, 10, 3000, 3064),
         (44121555, 0, 9, 10, 10830, 10840)]


def test_noalloc_equals_search_span():
    lib = gpu.load()
    with gpu.Context(0) as ctx:
        for span in SPANS:
            rc0, want, st0 = _call(lib, "ap26_search_span", ctx, span)
            rc1, got, st1 = _call(lib, "ap26_search_span_noalloc", ctx, span)
            assert rc0 == rc1 == gpu.AP26_OK, span
            assert got == want, span
            for f in ("k_searched", "survivors", "prp_tests", "reports", "positions", "walk_past_limit"):
                assert getattr(st1, f) == getattr(st0, f), (span, f)
        # again, smaller, straight after: nothing of the previous call leaks
        rc, again, _ = _call(lib, "ap26_search_span_noalloc", ctx, SPANS[1])
        rc0, want, _ = _call(lib, "ap26_search_span", ctx, SPANS[1])
        assert rc == rc0 == gpu.AP26_OK and again == want


def test_noalloc_limits():
    lib = gpu.load()
    with gpu.Context(0) as ctx:
        rc, _, _ = _call(lib, "ap26_search_span_noalloc", ctx, (366384, 0, 0, 10, 0, 65))
        assert rc == gpu.AP26_E_RANGE  # more n43 starts than the scratch is sized for
        rc, got, st = _call(lib, "ap26_search_span_noalloc", ctx, (366384, 0, 4, 4, 0, 64))
        assert rc == gpu.AP26_OK and got == [] and st.survivors == 0
        # capacity overflow reports the true count, as ap26_search_span does
        rc, got, st = _call(lib, "ap26_search_span_noalloc", ctx, (366384, 0, 0, 10, 0, 20), cap=1)
        assert rc == gpu.AP26_E_OVERFLOW and st.reports == 2
