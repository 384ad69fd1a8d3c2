"""ctypes mirror of the C ABI in include/ap26_gpu.h (libap26gpu.so).

This is the Python face of the search library: the same entry points, argument
meaning and error codes as the C header.  It loads the in-tree library and
raises if it is missing -- there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libap26gpu.so")

AP26_OK = 0
AP26_E_DEVICE = -1
AP26_E_NOMEM = -2
AP26_E_OVERFLOW = -3
AP26_E_RANGE = -4
AP26_E_INTERNAL = -5
AP26_E_BUSY = -6

NN43 = 10840
NSHIFT = 10


class Sol(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("K", ctypes.c_int32), ("first", ctypes.c_int64),
                ("range_id", ctypes.c_int32), ("shift_iter", ctypes.c_int32),
                ("prp2_mask", ctypes.c_uint32), ("valid_mask", ctypes.c_uint32)]


SOL_DTYPE = np.dtype([("k", "<i4"), ("K", "<i4"), ("first", "<i8"), ("range_id", "<i4"),
                      ("shift_iter", "<i4"), ("prp2_mask", "<u4"), ("valid_mask", "<u4")])
assert SOL_DTYPE.itemsize == ctypes.sizeof(Sol)


class Range(ctypes.Structure):
    _fields_ = [("kmin", ctypes.c_int32), ("kmax", ctypes.c_int32), ("shift", ctypes.c_int32),
                ("iter_start", ctypes.c_int32)]


class Opts(ctypes.Structure):
    _fields_ = [("sieve_variant", ctypes.c_int32), ("block", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("verbose", ctypes.c_int32)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("chunk", ctypes.c_int32), ("walk_depth", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("ms_per_k", ctypes.c_double), ("ms_walk", ctypes.c_double)]

    def as_dict(self):
        return {"variant": self.variant, "chunk": self.chunk, "walk_depth": self.walk_depth,
                "ms_per_k": self.ms_per_k, "ms_walk": self.ms_walk}


class Stats(ctypes.Structure):
    _fields_ = [("k_searched", ctypes.c_int64), ("k_skipped", ctypes.c_int64),
                ("positions", ctypes.c_int64), ("survivors", ctypes.c_int64),
                ("prp_tests", ctypes.c_int64), ("reports", ctypes.c_int64),
                ("ms_tables", ctypes.c_double), ("ms_sieve", ctypes.c_double),
                ("ms_walk", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("sieve_launches", ctypes.c_int32), ("walk_launches", ctypes.c_int32),
                ("walk_past_limit", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Every symbol the header declares, with its ctypes signature.
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
SIGNATURES = {
    "ap26_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "ap26_version": (ctypes.c_char_p, []),
    "ap26_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Opts), ctypes.POINTER(_P)]),
    "ap26_close": (ctypes.c_int, [_P]),
    "ap26_device_info": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I64),
                                        ctypes.POINTER(_I64), ctypes.c_char_p, _I32]),
    "ap26_search_k": (ctypes.c_int, [_P, _I32, _I32, _I32, _P, _I32, ctypes.POINTER(_I32),
                                     ctypes.POINTER(Stats)]),
    "ap26_search_span": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _I32,
                                        ctypes.POINTER(_I32), ctypes.POINTER(Stats)]),
    "ap26_search_span_noalloc": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _I32,
                                                ctypes.POINTER(_I32), ctypes.POINTER(Stats)]),
    "ap26_sieve_span": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _I64,
                                       ctypes.POINTER(_I64)]),
    "ap26_okok_table": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "ap26_batch_sieve_k": (ctypes.c_int, [_P, _I32, _I32, _I32, _P, _I64, ctypes.POINTER(_I64)]),
    "ap26_verify_sols": (ctypes.c_int, [_P, _P, _I32]),
    "ap26_prp_batch": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "ap26_prp_batch_dev": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_double)]),
    "ap26_search_batch": (ctypes.c_int, [_P, _P, _I32, _P, _I32, ctypes.POINTER(_I32), _P,
                                         ctypes.POINTER(Stats)]),
    "ap26_server_open": (ctypes.c_int, [_P, _I32, _I64, ctypes.POINTER(_P)]),
    "ap26_server_close": (ctypes.c_int, [_P]),
    "ap26_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Range), ctypes.POINTER(_I64)]),
    "ap26_server_pump": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "ap26_server_result": (ctypes.c_int, [_P, _I64, _P, _I32, ctypes.POINTER(_I32),
                                          ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(_I32)]),
    "ap26_server_progress": (ctypes.c_int, [_P, _I64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64),
                                            ctypes.POINTER(_I64)]),
    "ap26_server_stats": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                         ctypes.POINTER(_I64), ctypes.POINTER(_I32),
                                         ctypes.POINTER(_I32)]),
    "ap26_autotune": (ctypes.c_int, [_P, _I32, _I32, _I32, ctypes.c_char_p, ctypes.POINTER(TuneEntry),
                                     _I32, ctypes.POINTER(_I32), ctypes.POINTER(TuneEntry)]),
    "ap26_dev_alloc": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "ap26_dev_free": (ctypes.c_int, [_P, _P]),
    "ap26_copy_h2d": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "ap26_copy_d2h": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t]),
    "ap26_sync": (ctypes.c_int, [_P]),
}

_lib = None


def load(path=LIB_PATH):
    """Load libap26gpu.so (built in-tree by ap26_amd.build); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() "
                           "(there is no CPU fallback for the search)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Ap26Error(RuntimeError):
    """A negative AP26_E_* return; `stats` holds the call's ap26_stats when the
    entry point fills them on failure (AP26_E_RANGE from a walk that reached
    2^62 carries stats.walk_past_limit)."""

    def __init__(self, code, what, stats=None):
        self.code = code
        self.stats = stats
        super().__init__(f"{what}: {load().ap26_strerror(code).decode()} ({code})")


def _check(code, what, stats=None):
    if code != AP26_OK:
        raise Ap26Error(code, what, stats)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One ap26_ctx: device buffers and a stream on one GPU."""

    def __init__(self, device=0, sieve_variant=0, block=0, blocks_per_cu=0, verbose=0):
        lib = load()
        self._lib = lib
        opts = Opts(sieve_variant, block, blocks_per_cu, verbose)
        h = ctypes.c_void_p()
        _check(lib.ap26_open(device, ctypes.byref(opts), ctypes.byref(h)), "ap26_open")
        self.h = h

    def close(self):
        if self.h:
            self._lib.ap26_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        cu = ctypes.c_int32()
        lds = ctypes.c_int64()
        mem = ctypes.c_int64()
        name = ctypes.create_string_buffer(256)
        _check(self._lib.ap26_device_info(self.h, ctypes.byref(cu), ctypes.byref(lds),
                                          ctypes.byref(mem), name, 256), "ap26_device_info")
        return {"cu_count": cu.value, "lds_per_block": lds.value, "global_mem": mem.value,
                "arch": name.value.decode()}

    def search_k(self, K, shift, iter_start=0, cap=1 << 14):
        out = np.zeros(cap, dtype=SOL_DTYPE)
        n = ctypes.c_int32()
        st = Stats()
        rc = self._lib.ap26_search_k(self.h, K, shift, iter_start, _ptr(out), cap,
                                     ctypes.byref(n), ctypes.byref(st))
        _check(rc, "ap26_search_k", st)
        return out[:n.value].copy(), st

    def search_span(self, K, shift, it0, it1, j0, j1, cap=1 << 14):
        out = np.zeros(cap, dtype=SOL_DTYPE)
        n = ctypes.c_int32()
        st = Stats()
        rc = self._lib.ap26_search_span(self.h, K, shift, it0, it1, j0, j1, _ptr(out), cap,
                                        ctypes.byref(n), ctypes.byref(st))
        _check(rc, "ap26_search_span", st)
        return out[:n.value].copy(), st

    def search_span_noalloc(self, K, shift, it0, it1, j0, j1, cap=1 << 12):
        """ap26_search_span_noalloc: the tile search without any allocation
        (at most 64 n43 starts)."""
        out = np.zeros(cap, dtype=SOL_DTYPE)
        n = ctypes.c_int32()
        st = Stats()
        rc = self._lib.ap26_search_span_noalloc(self.h, K, shift, it0, it1, j0, j1, _ptr(out), cap,
                                                ctypes.byref(n), ctypes.byref(st))
        _check(rc, "ap26_search_span_noalloc", st)
        return out[:n.value].copy(), st

    def sieve_span(self, K, shift, it0, it1, j0, j1, cap=1 << 22):
        out = np.zeros(cap, dtype=np.int64)
        n = ctypes.c_int64()
        rc = self._lib.ap26_sieve_span(self.h, K, shift, it0, it1, j0, j1, _ptr(out), cap,
                                       ctypes.byref(n))
        _check(rc, "ap26_sieve_span")
        return out[:n.value].copy()

    def batch_sieve_k(self, K, shift, iter_start=0, cap=1 << 26):
        """Survivors of the batched engine's sieve for one K (parity against sieve_span)."""
        out = np.zeros(cap, dtype=np.int64)
        n = ctypes.c_int64()
        rc = self._lib.ap26_batch_sieve_k(self.h, K, shift, iter_start, _ptr(out), cap, ctypes.byref(n))
        _check(rc, "ap26_batch_sieve_k")
        return out[:n.value].copy()

    def verify_sols(self, sols):
        """Fill prp2_mask / valid_mask of a SOL_DTYPE array on the device."""
        sols = np.ascontiguousarray(sols, dtype=SOL_DTYPE)
        _check(self._lib.ap26_verify_sols(self.h, _ptr(sols), len(sols)), "ap26_verify_sols")
        return sols

    def okok_table(self, K, shift):
        out = np.zeros(23693, dtype=np.uint64)
        _check(self._lib.ap26_okok_table(self.h, K, shift, _ptr(out)), "ap26_okok_table")
        return out

    def prp_batch(self, n):
        n = np.ascontiguousarray(n, dtype=np.uint64)
        out = np.zeros(n.size, dtype=np.uint8)
        _check(self._lib.ap26_prp_batch(self.h, _ptr(n), _ptr(out), n.size), "ap26_prp_batch")
        return out

    def search_batch(self, ranges, cap=1 << 16):
        r = np.array([tuple(x) for x in ranges],
                     dtype=[("kmin", "<i4"), ("kmax", "<i4"), ("shift", "<i4"), ("iter_start", "<i4")])
        out = np.zeros(cap, dtype=SOL_DTYPE)
        n = ctypes.c_int32()
        tot = np.zeros(len(r), dtype=np.int64)
        st = Stats()
        rc = self._lib.ap26_search_batch(self.h, _ptr(r), len(r), _ptr(out), cap, ctypes.byref(n),
                                         _ptr(tot), ctypes.byref(st))
        _check(rc, "ap26_search_batch", st)
        return out[:n.value].copy(), tot, st

    def autotune(self, K=48620637, shift=640, quick=True, cache_path=None):
        """Pick the batched sieve variant / work-item size, then the walk's
        compaction depth, by measurement (ap26_autotune)."""
        res = (TuneEntry * 64)()
        n = ctypes.c_int32()
        best = TuneEntry()
        _check(self._lib.ap26_autotune(self.h, K, shift, 1 if quick else 0,
                                       cache_path.encode() if cache_path else None, res, 64,
                                       ctypes.byref(n), ctypes.byref(best)), "ap26_autotune")
        return best.as_dict(), [r.as_dict() for r in res[:n.value]]

    # device buffers
    def dev_alloc(self, nbytes):
        p = ctypes.c_void_p()
        _check(self._lib.ap26_dev_alloc(self.h, nbytes, ctypes.byref(p)), "ap26_dev_alloc")
        return p

    def dev_free(self, p):
        _check(self._lib.ap26_dev_free(self.h, p), "ap26_dev_free")

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        _check(self._lib.ap26_copy_h2d(self.h, dst, _ptr(arr), arr.nbytes), "ap26_copy_h2d")

    def d2h(self, arr, src):
        _check(self._lib.ap26_copy_d2h(self.h, _ptr(arr), src, arr.nbytes), "ap26_copy_d2h")

    def prp_batch_dev(self, d_n, d_out, count):
        ms = ctypes.c_double()
        _check(self._lib.ap26_prp_batch_dev(self.h, d_n, d_out, count, ctypes.byref(ms)),
               "ap26_prp_batch_dev")
        return ms.value

    def sync(self):
        _check(self._lib.ap26_sync(self.h), "ap26_sync")


class Server:
    """Continuous-batching server on a Context (ap26_server_* in the C ABI)."""

    def __init__(self, ctx, max_units=256, mem_budget_bytes=0):
        self._lib = ctx._lib
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(self._lib.ap26_server_open(ctx.h, max_units, mem_budget_bytes, ctypes.byref(h)),
               "ap26_server_open")
        self.h = h

    def close(self):
        if self.h:
            self._lib.ap26_server_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, kmin, kmax, shift, iter_start=0):
        """Returns a ticket, or None when the server is full (AP26_E_BUSY)."""
        r = Range(kmin, kmax, shift, iter_start)
        t = ctypes.c_int64()
        rc = self._lib.ap26_server_submit(self.h, ctypes.byref(r), ctypes.byref(t))
        if rc == AP26_E_BUSY:
            return None
        _check(rc, "ap26_server_submit")
        return t.value

    def pump(self):
        n = ctypes.c_int32()
        _check(self._lib.ap26_server_pump(self.h, ctypes.byref(n)), "ap26_server_pump")
        return n.value

    def result(self, ticket, cap=1 << 16):
        """None while running, else (reports, hash_total, latency_ms); a unit
        whose walk reached 2^62 raises Ap26Error(AP26_E_RANGE) once done."""
        out = np.zeros(cap, dtype=SOL_DTYPE)
        n = ctypes.c_int32()
        h = ctypes.c_int64()
        lat = ctypes.c_double()
        done = ctypes.c_int32()
        rc = self._lib.ap26_server_result(self.h, ticket, _ptr(out), cap, ctypes.byref(n),
                                          ctypes.byref(h), ctypes.byref(lat), ctypes.byref(done))
        _check(rc, "ap26_server_result")
        if not done.value:
            return None
        return out[:n.value].copy(), h.value, lat.value

    def progress(self, ticket):
        """(fraction done, searched K done, searched K in the unit)."""
        f, kd, kt = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.ap26_server_progress(self.h, ticket, ctypes.byref(f), ctypes.byref(kd), ctypes.byref(kt)),
               "ap26_server_progress")
        return f.value, kd.value, kt.value

    def stats(self):
        ud, kd, rd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        kpr, act = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.ap26_server_stats(self.h, ctypes.byref(ud), ctypes.byref(kd), ctypes.byref(rd),
                                           ctypes.byref(kpr), ctypes.byref(act)), "ap26_server_stats")
        return {"units_done": ud.value, "k_done": kd.value, "rounds": rd.value,
                "k_per_round": kpr.value, "active": act.value}
