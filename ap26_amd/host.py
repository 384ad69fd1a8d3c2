"""ctypes mirror of include/ap26_host.h (libap26host.so): the host side of the
AP26 operator surface -- state files, ReportSolution, result hash -- as the
AP26 command uses it.  CPU only.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libap26host.so")

AP26H_OK = 0
AP26H_E_IO = -10
AP26H_E_INVALID = -11
AP26H_E_RANGE = -12


class HostState(ctypes.Structure):
    _fields_ = [("KMIN", ctypes.c_int32), ("KMAX", ctypes.c_int32), ("SHIFT", ctypes.c_int32),
                ("K_COUNT", ctypes.c_int32), ("K_DONE", ctypes.c_int32),
                ("result_hash", ctypes.c_int32), ("last_trickle", ctypes.c_double),
                ("write_state_a_next", ctypes.c_int32), ("min_report", ctypes.c_int32),
                ("quiet", ctypes.c_int32), ("n_reports", ctypes.c_int64),
                ("n_records", ctypes.c_int64), ("n_nonsol", ctypes.c_int64),
                ("results_file", ctypes.c_void_p), ("state_a", ctypes.c_char * 1024),
                ("state_b", ctypes.c_char * 1024), ("results_name", ctypes.c_char * 1024),
                ("metrics_file", ctypes.c_void_p), ("t0_wall", ctypes.c_double), ("t0_cpu", ctypes.c_double)]


_P = ctypes.POINTER(HostState)
_I32 = ctypes.c_int32
SIGNATURES = {
    "ap26h_init": (None, [_P, _I32, _I32, _I32, ctypes.c_char_p, _I32]),
    "ap26h_close": (None, [_P]),
    "ap26h_will_search": (ctypes.c_int, [_I32]),
    "ap26h_count": (None, [_P, _I32]),
    "ap26h_read_state": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "ap26h_write_state": (None, [_P, _I32, _I32]),
    "ap26h_checkpoint": (None, [_P, _I32, _I32, _I32]),
    "ap26h_report": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_int64]),
    "ap26h_report_masks": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint32]),
    "ap26h_write_hash": (ctypes.c_int, [_P]),
    "ap26h_progress": (ctypes.c_double, [_P, _I32]),
    "ap26h_open_metrics": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "ap26h_hash_word": (ctypes.c_uint64, [_I32, _I32, _I32]),
    "ap26h_cap_total": (_I32, [ctypes.c_int64]),
    "ap26h_sprp": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64]),
    "ap26h_validate": (_I32, [_I32, _I32, ctypes.c_int64, _I32, ctypes.POINTER(_I32)]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class Host:
    """One work unit's host state (AP26 KMIN KMAX SHIFT) in directory `dir`."""

    def __init__(self, kmin, kmax, shift, dir="", min_report=20, quiet=True):
        self._lib = load()
        self.st = HostState()
        self._lib.ap26h_init(ctypes.byref(self.st), kmin, kmax, shift, dir.encode(), min_report)
        self.st.quiet = 1 if quiet else 0

    @property
    def hash(self):
        return self.st.result_hash

    def read_state(self):
        K = ctypes.c_int32()
        it = ctypes.c_int32()
        ok = self._lib.ap26h_read_state(ctypes.byref(self.st), ctypes.byref(K), ctypes.byref(it))
        return (K.value, it.value) if ok else None

    def write_state(self, K, it):
        self._lib.ap26h_write_state(ctypes.byref(self.st), K, it)

    def checkpoint(self, K, it, force=0):
        self._lib.ap26h_checkpoint(ctypes.byref(self.st), K, force, it)

    def report(self, k, K, first, prp2_mask=None, valid_mask=None):
        """ReportSolution; with the device's per-term masks (ap26_sol) the
        9-base verification is read from them instead of recomputed."""
        if prp2_mask is None:
            return self._lib.ap26h_report(ctypes.byref(self.st), k, K, first)
        return self._lib.ap26h_report_masks(ctypes.byref(self.st), k, K, first, prp2_mask, valid_mask)

    def progress(self, it=0):
        return self._lib.ap26h_progress(ctypes.byref(self.st), it)

    def open_metrics(self, path):
        return self._lib.ap26h_open_metrics(ctypes.byref(self.st), str(path).encode())

    def write_hash(self):
        return self._lib.ap26h_write_hash(ctypes.byref(self.st))

    def count(self, K_resume):
        self._lib.ap26h_count(ctypes.byref(self.st), K_resume)
        return self.st.K_COUNT, self.st.K_DONE

    def close(self):
        self._lib.ap26h_close(ctypes.byref(self.st))


def hash_word(kmin, kmax, h):
    return load().ap26h_hash_word(kmin, kmax, h)


def hash_line(kmin, kmax, h):
    return "%016X" % hash_word(kmin, kmax, h)


def cap_total(T):
    return load().ap26h_cap_total(T)


def will_search(K):
    return bool(load().ap26h_will_search(K))


def sprp(a, n):
    return bool(load().ap26h_sprp(a, n))


def validate(k, K, first, nbases=9):
    err = ctypes.c_int32()
    j = load().ap26h_validate(k, K, first, nbases, ctypes.byref(err))
    return j, err.value
