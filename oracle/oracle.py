"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference search (ap26_oracle.c).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
it, and only as the checker / timed CPU comparator.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

NPRIMES = 83
NOK = 23693
NN43 = 10840

SOL_DTYPE = np.dtype([("k", "<i4"), ("K", "<i4"), ("first", "<i8")])

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        sig = {
            "ap26o_will_search": (ctypes.c_int, [I32]),
            "ap26o_kconsts": (None, [I32, P]),
            "ap26o_n43_table": (None, [P, P]),
            "ap26o_primes": (ctypes.POINTER(I32), []),
            "ap26o_offsets": (ctypes.POINTER(I32), []),
            "ap26o_ok_table": (None, [I32, P]),
            "ap26o_okok_table": (None, [I32, I32, P]),
            "ap26o_sprp": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64]),
            "ap26o_prp2": (ctypes.c_int, [ctypes.c_uint64]),
            "ap26o_prp2_array": (None, [P, P, ctypes.c_size_t, I32]),
            "ap26o_walk": (I32, [I64, I64, ctypes.POINTER(I64)]),
            "ap26o_sieve_span": (I64, [I32, I32, I32, I32, I32, I32, P, I64]),
            "ap26o_search_span": (I64, [I32, I32, I32, I32, I32, I32, P, I64,
                                        ctypes.POINTER(I64), I32]),
            "ap26o_report": (ctypes.c_int, [I32, I32, I64, ctypes.POINTER(I32), I32, P, I64,
                                            ctypes.POINTER(I64), ctypes.POINTER(I64)]),
            "ap26o_validate": (I32, [I32, I32, I64, ctypes.c_int]),
            "ap26o_hash_word": (ctypes.c_uint64, [I32, I32, I32]),
            "ap26o_lookup_count": (I64, [I32, I32, I32, I32, I32, I32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class KConst(ctypes.Structure):
    _fields_ = [("K", ctypes.c_int32), ("STEP", ctypes.c_int64), ("n0", ctypes.c_int64),
                ("S31", ctypes.c_int64), ("S37", ctypes.c_int64), ("S41", ctypes.c_int64),
                ("S43", ctypes.c_int64), ("S47", ctypes.c_int64), ("S53", ctypes.c_int64),
                ("S59", ctypes.c_int64)]


def will_search(K):
    return bool(load().ap26o_will_search(K))


def kconsts(K):
    kc = KConst()
    load().ap26o_kconsts(K, ctypes.byref(kc))
    return kc


def n43_table(K):
    kc = kconsts(K)
    out = np.zeros(NN43, dtype=np.int64)
    load().ap26o_n43_table(ctypes.byref(kc), _ptr(out))
    return out


def primes():
    p = load().ap26o_primes()
    return [p[i] for i in range(NPRIMES)]


def offsets():
    p = load().ap26o_offsets()
    return [p[i] for i in range(NPRIMES)]


def ok_table(K):
    out = np.zeros(NOK, dtype=np.uint8)
    load().ap26o_ok_table(K, _ptr(out))
    return out


def okok_table(K, shift):
    out = np.zeros(NOK, dtype=np.uint64)
    load().ap26o_okok_table(K, shift, _ptr(out))
    return out


def sprp(base, n):
    return bool(load().ap26o_sprp(base, n))


def prp2(n):
    return bool(load().ap26o_prp2(n))


def prp2_array(arr, threads=1):
    arr = np.ascontiguousarray(arr, dtype=np.uint64)
    out = np.zeros(arr.size, dtype=np.uint8)
    load().ap26o_prp2_array(_ptr(arr), _ptr(out), ctypes.c_size_t(arr.size), threads)
    return out


def walk(n, step):
    first = ctypes.c_int64()
    k = load().ap26o_walk(n, step, ctypes.byref(first))
    return k, first.value


def sieve_span(K, shift, it0, it1, j0, j1, cap=1 << 22):
    out = np.zeros(cap, dtype=np.int64)
    n = load().ap26o_sieve_span(K, shift, it0, it1, j0, j1, _ptr(out), cap)
    if n > cap:
        raise RuntimeError("oracle survivor capacity exceeded")
    return out[:n].copy()


def search_span(K, shift, it0, it1, j0, j1, threads=1, cap=1 << 16):
    out = np.zeros(cap, dtype=SOL_DTYPE)
    ns = ctypes.c_int64()
    n = load().ap26o_search_span(K, shift, it0, it1, j0, j1, _ptr(out), cap, ctypes.byref(ns),
                                 threads)
    return out[:min(n, cap)].copy(), ns.value


def report(k, K, first, hash_value, min_report=20, cap=1 << 12):
    recs = np.zeros(cap, dtype=SOL_DTYPE)
    h = ctypes.c_int32(hash_value)
    nrec = ctypes.c_int64()
    nonsol = ctypes.c_int64()
    rc = load().ap26o_report(k, K, first, ctypes.byref(h), min_report, _ptr(recs), cap,
                             ctypes.byref(nrec), ctypes.byref(nonsol))
    return rc, h.value, recs[:nrec.value].copy(), nonsol.value


def validate(k, K, first, nbases=9):
    return load().ap26o_validate(k, K, first, nbases)


def hash_word(kmin, kmax, h):
    return load().ap26o_hash_word(kmin, kmax, h)


def lookup_count(K, shift, it0, it1, j0, j1):
    return load().ap26o_lookup_count(K, shift, it0, it1, j0, j1)
