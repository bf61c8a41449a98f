"""ctypes wrapper of the CPU oracle (oracle/hb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's CPU-baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import Callable, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

HASH_SIPHASH13, HASH_JENKINS = 0, 1
OR_MAX_WORDS = 6144  # hb_oracle.h


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class _Hll(ctypes.Structure):
    _fields_ = [("log2m", ctypes.c_uint32), ("m", ctypes.c_uint32), ("r", ctypes.c_uint32),
                ("W", ctypes.c_uint32), ("sentinel", ctypes.c_uint64), ("alpha_m_m", ctypes.c_double),
                ("hash", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("msb", ctypes.c_uint64 * OR_MAX_WORDS), ("lsb", ctypes.c_uint64 * OR_MAX_WORDS)]


_P, _U64, _U32, _D, _I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_int
DISCOUNT_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p)
_SIGS = [
    ("or_register_size", _U32, [_U64]),
    ("or_rel_std", _D, [_U32]),
    ("or_log2_num_registers", _U32, [_D]),
    ("or_hll_init", _I, [_P, _U32, _U32, _U64, _I, _U64]),
    ("or_jenkins", _U64, [_U64, _U64]),
    ("or_siphash", _U64, [_P, ctypes.c_size_t, _U64, _U64, _I, _I]),
    ("or_siphash13_u64", _U64, [_U64]),
    ("or_hash_item", _U64, [_P, _U64]),
    ("or_weight_item", _U64, [_U64, _U64]),
    ("or_get_register", _U64, [_P, _P, _U32]),
    ("or_set_register", None, [_P, _P, _U32, _U64]),
    ("or_add", None, [_P, _P, _U64]),
    ("or_count", _D, [_P, _P]),
    ("or_merge_broadword", None, [_P, _P, _P]),
    ("or_merge_max", None, [_P, _P, _P]),
    ("or_jacobi_init", None, [_P, _U64, _P, _P, _U64]),
    ("or_jacobi_iterate", _U64, [_P, _U64, _P, _P, _P, _P, _P]),
    ("or_hb_create", _P, [_U64, _U64, _P, _P, _P, _P, _U32, _U32, _I, _U64, _I, _I, _P, _U64, _I]),
    ("or_hb_add_discount", _I, [_P, DISCOUNT_FN, _P]),
    ("or_hb_destroy", None, [_P]),
    ("or_hb_init", None, [_P]),
    ("or_hb_iterate", None, [_P]),
    ("or_hb_run", _I, [_P, _U64, _D]),
    ("or_hb_iteration", _U64, [_P]),
    ("or_hb_modified", _U64, [_P]),
    ("or_hb_visited_arcs", _U64, [_P]),
    ("or_hb_relative_increment", _D, [_P]),
    ("or_hb_mode", _I, [_P]),
    ("or_hb_nf", ctypes.c_size_t, [_P, _P, ctypes.c_size_t]),
    ("or_hb_counters", ctypes.POINTER(ctypes.c_uint64), [_P, _I]),
    ("or_hb_modified_bits", ctypes.POINTER(ctypes.c_uint64), [_P]),
    ("or_hb_sum_of_distances", ctypes.POINTER(ctypes.c_double), [_P]),
    ("or_hb_harmonic", ctypes.POINTER(ctypes.c_double), [_P]),
    ("or_hb_discounted", ctypes.POINTER(ctypes.c_double), [_P, ctypes.c_size_t]),
    ("or_hb_derived", _I, [_P, _I, _P]),
]

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


class Hll:
    """HyperLogLog logic (hyper_log_log_logic.rs:374-428) over one counter."""

    def __init__(self, log2m: int, reg_bits: int = 0, num_elements: int = 0,
                 hash: int = HASH_SIPHASH13, seed: int = 0):
        self.h = _Hll()
        rc = lib().or_hll_init(ctypes.byref(self.h), log2m, reg_bits, num_elements, hash, seed)
        if rc != 0:
            raise ValueError(f"or_hll_init failed with {rc}")

    @property
    def W(self) -> int:
        return self.h.W

    @property
    def m(self) -> int:
        return self.h.m

    @property
    def r(self) -> int:
        return self.h.r

    def new_counter(self) -> np.ndarray:
        return np.zeros(self.h.W, dtype=np.uint64)

    def get(self, c: np.ndarray, i: int) -> int:
        return lib().or_get_register(ctypes.byref(self.h), _p(c), i)

    def set(self, c: np.ndarray, i: int, v: int) -> None:
        lib().or_set_register(ctypes.byref(self.h), _p(c), i, v)

    def add(self, c: np.ndarray, item: int) -> None:
        lib().or_add(ctypes.byref(self.h), _p(c), item)

    def count(self, c: np.ndarray) -> float:
        return lib().or_count(ctypes.byref(self.h), _p(c))

    def merge_broadword(self, x: np.ndarray, y: np.ndarray) -> None:
        lib().or_merge_broadword(ctypes.byref(self.h), _p(x), _p(y))

    def merge_max(self, x: np.ndarray, y: np.ndarray) -> None:
        lib().or_merge_max(ctypes.byref(self.h), _p(x), _p(y))

    def registers(self, c: np.ndarray) -> np.ndarray:
        return np.array([self.get(c, i) for i in range(self.h.m)], dtype=np.uint64)

    def counts(self, counters: np.ndarray) -> np.ndarray:
        counters = np.ascontiguousarray(counters, dtype=np.uint64).reshape(-1, self.h.W)
        return np.array([self.count(row) for row in counters])


class Jacobi:
    """Sequential Jacobi HyperBall (hyperball_impl.rs:1230-1261)."""

    def __init__(self, hll: Hll, offsets: np.ndarray, successors: np.ndarray,
                 weights: Optional[np.ndarray] = None, weight_seed: int = 0):
        self.hll = hll
        self.off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.succ = np.ascontiguousarray(successors, dtype=np.uint32)
        self.n = len(self.off) - 1
        self.curr = np.zeros((self.n, hll.W), dtype=np.uint64)
        self.next = np.zeros_like(self.curr)
        self.modified = np.zeros((self.n + 63) // 64, dtype=np.uint64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint64)
        lib().or_jacobi_init(ctypes.byref(hll.h), self.n, _p(self.curr), _p(w) if w is not None else None,
                             weight_seed)

    def iterate(self) -> int:
        k = lib().or_jacobi_iterate(ctypes.byref(self.hll.h), self.n, _p(self.off), _p(self.succ),
                                    _p(self.curr), _p(self.next), _p(self.modified))
        self.curr, self.next = self.next, self.curr
        return int(k)


class OracleHyperBall:
    """Restatement of the reference engine (hyperball_impl.rs:534-1213)."""

    def __init__(self, offsets, successors, t_offsets=None, t_successors=None, *, log2m: int,
                 reg_bits: int = 0, hash: int = HASH_SIPHASH13, seed: int = 0,
                 sum_of_distances: bool = False, sum_of_inverse_distances: bool = False,
                 weights=None, weight_seed: int = 0, threads: int = 0,
                 discount_functions: tuple = ()):
        self.off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.succ = np.ascontiguousarray(successors, dtype=np.uint32)
        self.toff = None if t_offsets is None else np.ascontiguousarray(t_offsets, dtype=np.uint64)
        self.tsucc = None if t_successors is None else np.ascontiguousarray(t_successors, dtype=np.uint32)
        self.w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint64)
        self.n = len(self.off) - 1
        self.hll = Hll(log2m, reg_bits, int(self.w.sum()) if self.w is not None else self.n, hash, seed)
        self.h = lib().or_hb_create(
            self.n, len(self.succ), _p(self.off), _p(self.succ),
            _p(self.toff) if self.toff is not None else None,
            _p(self.tsucc) if self.tsucc is not None else None,
            log2m, reg_bits, hash, seed, int(sum_of_distances), int(sum_of_inverse_distances),
            _p(self.w) if self.w is not None else None, weight_seed, threads)
        if not self.h:
            raise ValueError("or_hb_create failed")
        self._cbs = []
        for f in discount_functions:
            cb = DISCOUNT_FN(lambda d, _c, f=f: float(f(int(d))))
            self._cbs.append(cb)
            lib().or_hb_add_discount(self.h, cb, None)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_hb_destroy(self.h)
            self.h = None

    def run(self, upper_bound: int = 2**64 - 1, threshold: Optional[float] = None) -> int:
        return lib().or_hb_run(self.h, upper_bound, math.nan if threshold is None else threshold)

    def init(self) -> None:
        lib().or_hb_init(self.h)

    def iterate(self) -> None:
        lib().or_hb_iterate(self.h)

    @property
    def iteration(self) -> int:
        return lib().or_hb_iteration(self.h)

    @property
    def modified(self) -> int:
        return lib().or_hb_modified(self.h)

    @property
    def visited_arcs(self) -> int:
        return lib().or_hb_visited_arcs(self.h)

    @property
    def mode(self) -> int:
        return lib().or_hb_mode(self.h)

    def neighborhood_function(self) -> np.ndarray:
        k = lib().or_hb_nf(self.h, None, 0)
        out = np.zeros(k)
        lib().or_hb_nf(self.h, _p(out), k)
        return out

    def counters(self, which: int = 0) -> np.ndarray:
        p = lib().or_hb_counters(self.h, which)
        return np.ctypeslib.as_array(p, (self.n, self.hll.W)).copy()

    def modified_bits(self) -> np.ndarray:
        p = lib().or_hb_modified_bits(self.h)
        return np.ctypeslib.as_array(p, ((self.n + 63) // 64,)).copy()

    def sum_of_distances(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_hb_sum_of_distances(self.h), (self.n,)).copy()

    def harmonic_centralities(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_hb_harmonic(self.h), (self.n,)).copy()

    def discounted_centrality(self, k: int) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_hb_discounted(self.h, k), (self.n,)).copy()

    def _derived(self, which: int) -> np.ndarray:
        out = np.zeros(self.n)
        if lib().or_hb_derived(self.h, which, _p(out)) != 0:
            raise ValueError("sum of distances not requested")
        return out

    def closeness_centrality(self) -> np.ndarray:
        return self._derived(0)

    def lin_centrality(self) -> np.ndarray:
        return self._derived(1)

    def nieminen_centrality(self) -> np.ndarray:
        return self._derived(2)

    def reachable_nodes(self) -> np.ndarray:
        return self._derived(3)


def jenkins(x: int, seed: int) -> int:
    return lib().or_jenkins(x, seed)


def siphash(msg: bytes, k0: int, k1: int, c: int, d: int) -> int:
    buf = ctypes.create_string_buffer(msg, len(msg))
    return lib().or_siphash(buf, len(msg), k0, k1, c, d)


def siphash13_u64(x: int) -> int:
    return lib().or_siphash13_u64(x)


def register_size(n: int) -> int:
    return lib().or_register_size(n)


def rel_std(log2m: int) -> float:
    return lib().or_rel_std(log2m)


def log2_num_registers(rsd: float) -> int:
    return lib().or_log2_num_registers(rsd)
