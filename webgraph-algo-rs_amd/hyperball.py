"""Python mirror of the reference crate's HyperBall API over the C ABI of
libhyperball_amd.so (include/hyperball.h).

Names and argument meaning follow the reference (vigna/webgraph-algo-rs,
src/algo/hyperball/hyperball_impl.rs):

    hb = (HyperBallBuilder.with_hyper_log_log(graph, transpose, log2m=8)
          .sum_of_distances(True).sum_of_inverse_distances(True)
          .build())
    hb.run_until_done()
    hb.harmonic_centralities()

Errors the reference reports through anyhow::Result (or panics) raise
HyperBallError with the status code of the C ABI and the reference's message.
There is no CPU fallback: without the built library or a gfx950 device the
calls fail.
"""
from __future__ import annotations

import ctypes
import math
import os
import weakref
from typing import Callable, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhyperball_amd.so")

HB_OK, HB_E_NOT_RUN, HB_E_NOT_REQUESTED, HB_E_INVALID_ARG, HB_E_OOM, HB_E_HIP, HB_E_NO_DEVICE = range(7)
HASH_SIPHASH13, HASH_JENKINS = 0, 1
ACC_F64, ACC_F32_KAHAN, ACC_F32 = 0, 1, 2
COUNTERS_DEVICE, COUNTERS_HOST, COUNTERS_STREAM = 0, 1, 2


class HyperBallError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class _Csr(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint64), ("num_arcs", ctypes.c_uint64),
                ("offsets", ctypes.c_void_p), ("successors", ctypes.c_void_p)]


class _Graph(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint64), ("num_arcs", ctypes.c_uint64),
                ("offsets", ctypes.POINTER(ctypes.c_uint64)),
                ("successors", ctypes.POINTER(ctypes.c_uint32))]


class _BvParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in (
        "window_size", "max_ref_count", "min_interval_length", "zeta_k", "outdegrees",
        "references", "blocks", "block_count", "residuals", "offsets")]


CODE_GAMMA, CODE_DELTA, CODE_UNARY, CODE_ZETA = 0, 1, 2, 3


class _Config(ctypes.Structure):
    _fields_ = [("log2m", ctypes.c_uint32), ("reg_bits", ctypes.c_uint32), ("hash", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("num_elements", ctypes.c_uint64),
                ("weights", ctypes.c_void_p), ("weight_seed", ctypes.c_uint64),
                ("arc_granularity", ctypes.c_uint64), ("sum_of_distances", ctypes.c_int),
                ("sum_of_inverse_distances", ctypes.c_int), ("acc_precision", ctypes.c_uint32),
                ("device", ctypes.c_int), ("device_transpose", ctypes.c_int),
                ("counter_memory", ctypes.c_uint32), ("stream_tile_nodes", ctypes.c_uint64)]


class _Params(ctypes.Structure):
    _fields_ = [("log2m", ctypes.c_uint32), ("num_registers", ctypes.c_uint32),
                ("register_bits", ctypes.c_uint32), ("words_per_counter", ctypes.c_uint32),
                ("num_nodes", ctypes.c_uint64), ("num_arcs", ctypes.c_uint64),
                ("alpha_m_m", ctypes.c_double), ("counter_memory", ctypes.c_uint32)]


class IterationStats(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_uint64), ("systolic", ctypes.c_uint32),
                ("local", ctypes.c_uint32), ("pre_local", ctypes.c_uint32),
                ("checked", ctypes.c_uint64), ("visited_arcs", ctypes.c_uint64),
                ("scanned_arcs", ctypes.c_uint64), ("transpose_arcs", ctypes.c_uint64),
                ("copied", ctypes.c_uint64), ("modified", ctypes.c_uint64), ("nf", ctypes.c_double),
                ("union_ms", ctypes.c_double), ("iteration_ms", ctypes.c_double),
                ("gathered_slabs", ctypes.c_uint64), ("delta_gathering", ctypes.c_uint32),
                ("streamed", ctypes.c_uint32), ("window", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


DISCOUNT_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(IterationStats), ctypes.c_void_p)

# (name, restype, argtypes) of every entry point of include/hyperball.h
_P, _U64, _U32, _D, _I, _SZ = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
_SIGNATURES = [
    ("hb_last_error", ctypes.c_char_p, []),
    ("hb_bvgraph_load", _I, [ctypes.c_char_p, _P]),
    ("hb_bvgraph_load_offsets", _I, [ctypes.c_char_p, _U64, _P]),
    ("hb_bvgraph_decoder_offsets", _I, [ctypes.c_char_p, _P, _U64]),
    ("hb_elias_fano_load", _I, [ctypes.c_char_p, _P, _U64, ctypes.POINTER(ctypes.c_uint64)]),
    ("hb_graph_transpose", _I, [_P, _P]),
    ("hb_bvgraph_load_threads", _I, [_P, _U32, _P]),
    ("hb_bvgraph_params_init", None, [_P]),
    ("hb_bvgraph_store", _I, [_P, ctypes.c_char_p, _P]),
    ("hb_graph_transpose_device", _I, [_P, _I, _P]),
    ("hb_graph_generate_uniform", _I, [_U64, _U32, _U64, _P]),
    ("hb_graph_generate_web", _I, [_U64, _D, _U32, _U32, _U64, _D, _U64, _P]),
    ("hb_graph_free", None, [_P]),
    ("hb_register_size", _U32, [_U64]),
    ("hb_rel_std", _D, [_U32]),
    ("hb_log2_num_registers", _U32, [_D]),
    ("hb_config_init", None, [_P]),
    ("hb_create", _I, [_P, _P, _P, _P]),
    ("hb_destroy", _I, [_P]),
    ("hb_add_discount_function", _I, [_P, DISCOUNT_FN, _P]),
    ("hb_clear_discount_functions", _I, [_P]),
    ("hb_run", _I, [_P, _U64, _D]),
    ("hb_run_until_stable", _I, [_P, _U64]),
    ("hb_run_until_done", _I, [_P]),
    ("hb_init", _I, [_P]),
    ("hb_iterate", _I, [_P]),
    ("hb_neighborhood_function", _I, [_P, _P, _SZ, _P]),
    ("hb_sum_of_distances", _I, [_P, _P, _SZ]),
    ("hb_harmonic_centralities", _I, [_P, _P, _SZ]),
    ("hb_discounted_centrality", _I, [_P, _SZ, _P, _SZ]),
    ("hb_closeness_centrality", _I, [_P, _P, _SZ]),
    ("hb_lin_centrality", _I, [_P, _P, _SZ]),
    ("hb_nieminen_centrality", _I, [_P, _P, _SZ]),
    ("hb_reachable_nodes", _I, [_P, _P, _SZ]),
    ("hb_reachable_nodes_from", _I, [_P, _U64, _P]),
    ("hb_get_params", _I, [_P, _P]),
    ("hb_get_counters", _I, [_P, _I, _P, _SZ]),
    ("hb_get_counters_noalloc", _I, [_P, _I, _P, _SZ]),
    ("hb_get_modified", _I, [_P, _P, _SZ]),
    ("hb_num_iterations", _I, [_P, _P]),
    ("hb_get_iteration_stats", _I, [_P, _U64, _P]),
    ("hb_set_timing", _I, [_P, _I]),
    ("hb_set_tuning", _I, [_P, _U32]),
    ("hb_batch_create", _I, [_SZ, _P, _P, _P, _P]),
    ("hb_batch_destroy", _I, [_P]),
    ("hb_batch_instance", _I, [_P, _SZ, _P]),
    ("hb_batch_run", _I, [_P, _U64, _D]),
    ("hb_batch_run_until_done", _I, [_P]),
    ("hb_batch_set_graph", _I, [_P, _I]),
    ("hb_autotune", _I, [_P, _U32, _P, _P, _SZ, _P, _P]),
    ("hb_set_delta_gathering", _I, [_P, _I]),
    ("hb_set_window", _I, [_P, _I, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("hb_set_streaming", _I, [_P, _I]),
    ("hb_set_progress", _I, [_P, _P, _P]),
    ("hb_set_stream", _I, [_P, _P]),
    ("hb_batch_autotune", _I, [_P, _U32, _P, _P, _SZ, _P, _P]),
    ("hb_batch_step_times", _I, [_P, _P, _SZ, _P]),
]

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded C ABI library (built by webgraph-algo-rs_amd/build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != HB_OK:
        raise HyperBallError(status, lib().hb_last_error().decode())


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ---------------------------------------------------------------------------
# HyperLogLog statics (hyper_log_log_logic.rs:275-303)
# ---------------------------------------------------------------------------

def register_size(n: int) -> int:
    return lib().hb_register_size(n)


def rel_std(log2m: int) -> float:
    return lib().hb_rel_std(log2m)


def log_2_num_of_registers(rsd: float) -> int:
    return lib().hb_log2_num_registers(rsd)


# ---------------------------------------------------------------------------
# Graphs
# ---------------------------------------------------------------------------

def load_elias_fano(path: str) -> np.ndarray:
    """The elements of an Elias-Fano sequence that the sux crate stored
    through epserde: a BVGraph's .ef (bit offsets of the node records) or
    .dcf (cumulative out-degrees)."""
    cnt = ctypes.c_uint64()
    _check(lib().hb_elias_fano_load(path.encode(), None, 0, ctypes.byref(cnt)))
    out = np.empty(cnt.value, dtype=np.uint64)
    _check(lib().hb_elias_fano_load(path.encode(), _ptr(out), out.size, ctypes.byref(cnt)))
    return out


class Graph:
    """A CSR graph: successors of v are successors[offsets[v]:offsets[v+1]]."""

    def __init__(self, offsets: np.ndarray, successors: np.ndarray, _owner=None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.successors = np.ascontiguousarray(successors, dtype=np.uint32)
        self._owner = _owner
        if self.offsets.ndim != 1 or len(self.offsets) < 1:
            raise ValueError("offsets must have num_nodes + 1 entries")

    @property
    def num_nodes(self) -> int:
        return len(self.offsets) - 1

    @property
    def num_arcs(self) -> int:
        return len(self.successors)

    def outdegrees(self) -> np.ndarray:
        return np.diff(self.offsets)

    def csr(self) -> _Csr:
        return _Csr(self.num_nodes, self.num_arcs, _ptr(self.offsets), _ptr(self.successors))

    @classmethod
    def _from_lib(cls, g: _Graph) -> "Graph":
        n, a = g.num_nodes, g.num_arcs
        off = np.ctypeslib.as_array(g.offsets, (n + 1,))
        succ = np.ctypeslib.as_array(g.successors, (max(a, 1),))[:a]
        holder = _GraphHolder(g)
        return cls(off, succ, _owner=holder)

    @classmethod
    def from_arcs(cls, n: int, arcs) -> "Graph":
        arcs = np.asarray(arcs, dtype=np.int64).reshape(-1, 2)
        if len(arcs):
            arcs = np.unique(arcs, axis=0)
        counts = np.bincount(arcs[:, 0], minlength=n) if len(arcs) else np.zeros(n, np.int64)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
        return cls(off, arcs[:, 1].astype(np.uint32) if len(arcs) else np.zeros(0, np.uint32))

    @classmethod
    def load_bvgraph(cls, basename: str, threads: Optional[int] = None) -> "Graph":
        """Decodes <basename>.graph (BVGraph v0); threads None = all cores,
        in parallel through the node index, <basename>.ef or .offsets, when
        one exists."""
        g = _Graph()
        if threads is None:
            _check(lib().hb_bvgraph_load(basename.encode(), ctypes.byref(g)))
        else:
            _check(lib().hb_bvgraph_load_threads(basename.encode(), int(threads), ctypes.byref(g)))
        return cls._from_lib(g)

    def store_bvgraph(self, basename: str, **params) -> None:
        """Writes the graph as a BVGraph (<basename>.graph, .offsets,
        .properties); params override the compressor defaults (window_size,
        max_ref_count, min_interval_length, zeta_k and the CODE_* of
        outdegrees, references, blocks, block_count, residuals, offsets)."""
        p = _BvParams()
        lib().hb_bvgraph_params_init(ctypes.byref(p))
        for k, v in params.items():
            if k not in dict(_BvParams._fields_):
                raise TypeError(f"unknown BVGraph parameter {k}")
            setattr(p, k, int(v))
        c = self.csr()
        _check(lib().hb_bvgraph_store(ctypes.byref(c), basename.encode(), ctypes.byref(p)))

    @classmethod
    def uniform(cls, n: int, k: int, seed: int) -> "Graph":
        g = _Graph()
        _check(lib().hb_graph_generate_uniform(n, k, seed, ctypes.byref(g)))
        return cls._from_lib(g)

    @classmethod
    def web(cls, n: int, alpha: float = 2.1, dmin: int = 3, dmax: int = 10000, window: int = 1000,
            p_window: float = 0.7, seed: int = 2) -> "Graph":
        g = _Graph()
        _check(lib().hb_graph_generate_web(n, alpha, dmin, dmax, window, p_window, seed, ctypes.byref(g)))
        return cls._from_lib(g)

    def transpose(self, device: Optional[int] = None) -> "Graph":
        """The transposed graph, built on the host (counting sort) or, with
        `device`, on that HIP device (identical arrays)."""
        g = _Graph()
        c = self.csr()
        if device is None:
            _check(lib().hb_graph_transpose(ctypes.byref(c), ctypes.byref(g)))
        else:
            _check(lib().hb_graph_transpose_device(ctypes.byref(c), int(device), ctypes.byref(g)))
        return Graph._from_lib(g)


class _GraphHolder:
    def __init__(self, g: _Graph):
        self.g = g
        weakref.finalize(self, lib().hb_graph_free, ctypes.byref(g))


# ---------------------------------------------------------------------------
# Builder and engine (hyperball_impl.rs:26-404, :477-757)
# ---------------------------------------------------------------------------

class HyperBallBuilder:
    """HyperBallBuilder (hyperball_impl.rs:26-56)."""

    DEFAULT_GRANULARITY = 16 * 1024  # hyperball_impl.rs:193

    def __init__(self, graph: Graph, transpose: Optional[Graph], log2m: int,
                 weights: Optional[np.ndarray] = None):
        self._graph = graph
        self._transpose = transpose
        self._log2m = log2m
        self._weights = None
        self._reg_bits = 0
        self._hash = HASH_SIPHASH13
        self._seed = 0
        self._sd = False
        self._hc = False
        self._granularity = self.DEFAULT_GRANULARITY
        self._discounts: list[Callable[[int], float]] = []
        self._acc = ACC_F64
        self._device = 0
        self._device_transpose = False
        self._counter_memory = COUNTERS_DEVICE
        self._stream_tile_nodes = 0
        self._weight_seed = 0
        self._num_elements = 0
        if weights is not None:
            self.weights(weights)

    @classmethod
    def with_hyper_log_log(cls, graph: Graph, transposed: Optional[Graph], log2m: int,
                           weights: Optional[np.ndarray] = None) -> "HyperBallBuilder":
        """HyperBallBuilder::with_hyper_log_log (hyperball_impl.rs:91-133); the
        cumulative outdegree function is the CSR offsets themselves."""
        if weights is not None and len(weights) != graph.num_nodes:
            raise HyperBallError(HB_E_INVALID_ARG,
                                 "weights should have length equal to the graph's number of nodes")
        return cls(graph, transposed, log2m, weights)

    @classmethod
    def with_transpose(cls, graph: Graph, transpose: Graph, log2m: int) -> "HyperBallBuilder":
        """HyperBallBuilder::with_transpose (hyperball_impl.rs:204-256)."""
        return cls(graph, transpose, log2m)

    # --- setters (hyperball_impl.rs:258-303) ---
    def sum_of_distances(self, b: bool) -> "HyperBallBuilder":
        self._sd = bool(b)
        return self

    def sum_of_inverse_distances(self, b: bool) -> "HyperBallBuilder":
        self._hc = bool(b)
        return self

    def arc_granularity(self, g: int) -> "HyperBallBuilder":
        self._granularity = int(g)
        return self

    def weights(self, w: Optional[np.ndarray]) -> "HyperBallBuilder":
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.uint64)
            if len(w) != self._graph.num_nodes:
                raise HyperBallError(HB_E_INVALID_ARG,
                                     f"assertion failed: weights length {len(w)} != {self._graph.num_nodes}")
        self._weights = w
        return self

    def discount_function(self, f: Callable[[int], float]) -> "HyperBallBuilder":
        self._discounts.append(f)
        return self

    def no_discount_function(self) -> "HyperBallBuilder":
        self._discounts.clear()
        return self

    # --- counter-logic options (HyperLogLogBuilder, hyper_log_log_logic.rs:306-363) ---
    def jenkins(self, seed: int) -> "HyperBallBuilder":
        """Use the Jenkins hasher of tests/test_hyperball.rs:26-172 with `seed`."""
        self._hash, self._seed = HASH_JENKINS, int(seed)
        return self

    def default_hasher(self) -> "HyperBallBuilder":
        """std DefaultHasher (SipHash-1-3 with zero keys), the reference default."""
        self._hash, self._seed = HASH_SIPHASH13, 0
        return self

    def register_bits(self, r: int) -> "HyperBallBuilder":
        """Register width override (0 = HyperLogLog::register_size)."""
        self._reg_bits = int(r)
        return self

    def num_elements(self, n: int) -> "HyperBallBuilder":
        self._num_elements = int(n)
        return self

    def weight_seed(self, s: int) -> "HyperBallBuilder":
        self._weight_seed = int(s)
        return self

    def transpose_on_device(self, on: bool = True) -> "HyperBallBuilder":
        """With no transpose given, build it on the device inside build()
        (enables the systolic mode without a host transpose)."""
        self._device_transpose = bool(on)
        return self

    def counter_memory(self, where: int, tile_nodes: int = 0) -> "HyperBallBuilder":
        """Where the two counter arrays live (the reference's mmap_options
        argument, hyperball_impl.rs:89-117): COUNTERS_DEVICE (HBM, default),
        COUNTERS_HOST (pinned host memory read and written by the kernels
        over the host link) or COUNTERS_STREAM (pinned host memory streamed
        through device memory in tiles of tile_nodes nodes, 0: automatic),
        for counters beyond HBM."""
        self._counter_memory = int(where)
        self._stream_tile_nodes = int(tile_nodes)
        return self

    def accumulator_precision(self, mode: int) -> "HyperBallBuilder":
        self._acc = int(mode)
        return self

    def device(self, d: int) -> "HyperBallBuilder":
        self._device = int(d)
        return self

    def _config(self) -> _Config:
        cfg = _Config()
        lib().hb_config_init(ctypes.byref(cfg))
        cfg.log2m = self._log2m
        cfg.reg_bits = self._reg_bits
        cfg.hash = self._hash
        cfg.seed = self._seed
        cfg.num_elements = self._num_elements
        cfg.weights = _ptr(self._weights) if self._weights is not None else None
        cfg.weight_seed = self._weight_seed
        cfg.arc_granularity = self._granularity
        cfg.sum_of_distances = int(self._sd)
        cfg.sum_of_inverse_distances = int(self._hc)
        cfg.acc_precision = self._acc
        cfg.device = self._device
        cfg.device_transpose = int(self._device_transpose)
        cfg.counter_memory = self._counter_memory
        cfg.stream_tile_nodes = self._stream_tile_nodes
        return cfg

    def build(self) -> "HyperBall":
        """HyperBallBuilder::build (hyperball_impl.rs:321-403)."""
        cfg = self._config()
        g = self._graph.csr()
        t = self._transpose.csr() if self._transpose is not None else None
        h = ctypes.c_void_p()
        _check(lib().hb_create(ctypes.byref(g), ctypes.byref(t) if t is not None else None,
                               ctypes.byref(cfg), ctypes.byref(h)))
        hb = HyperBall(h, self._graph.num_nodes)
        for f in self._discounts:
            hb._add_discount(f)
        return hb


class HyperBall:
    """HyperBall (hyperball_impl.rs:477-509) backed by a device instance."""

    def __init__(self, handle: ctypes.c_void_p, n: int, owner=None):
        self._h = handle
        self._n = n
        self._callbacks = []
        self._owner = owner  # a batch that owns the instance
        self._fin = weakref.finalize(self, lambda: None) if owner is not None else \
            weakref.finalize(self, lib().hb_destroy, handle)

    def close(self) -> None:
        self._fin()

    def _add_discount(self, f: Callable[[int], float]) -> None:
        cb = DISCOUNT_FN(lambda d, _ctx: float(f(int(d))))
        self._callbacks.append(cb)
        _check(lib().hb_add_discount_function(self._h, cb, None))

    # --- running (hyperball_impl.rs:534-621) ---
    def run(self, upper_bound: int, threshold: Optional[float] = None) -> None:
        ub = min(int(upper_bound), 2**64 - 1)
        _check(lib().hb_run(self._h, ub, math.nan if threshold is None else float(threshold)))

    def run_until_stable(self, upper_bound: int) -> None:
        _check(lib().hb_run_until_stable(self._h, min(int(upper_bound), 2**64 - 1)))

    def run_until_done(self) -> None:
        _check(lib().hb_run_until_done(self._h))

    def init(self) -> None:
        _check(lib().hb_init(self._h))

    def iterate(self) -> None:
        _check(lib().hb_iterate(self._h))

    # --- results (hyperball_impl.rs:633-757) ---
    def neighborhood_function(self) -> np.ndarray:
        ln = ctypes.c_size_t()
        _check(lib().hb_neighborhood_function(self._h, None, 0, ctypes.byref(ln)))
        out = np.zeros(ln.value, dtype=np.float64)
        _check(lib().hb_neighborhood_function(self._h, _ptr(out), len(out), ctypes.byref(ln)))
        return out

    def _vec(self, fn, *args) -> np.ndarray:
        out = np.zeros(self._n, dtype=np.float64)
        _check(fn(self._h, *args, _ptr(out), self._n))
        return out

    def sum_of_distances(self) -> np.ndarray:
        return self._vec(lib().hb_sum_of_distances)

    def harmonic_centralities(self) -> np.ndarray:
        return self._vec(lib().hb_harmonic_centralities)

    def discounted_centrality(self, index: int) -> np.ndarray:
        return self._vec(lib().hb_discounted_centrality, index)

    def closeness_centrality(self) -> np.ndarray:
        return self._vec(lib().hb_closeness_centrality)

    def lin_centrality(self) -> np.ndarray:
        return self._vec(lib().hb_lin_centrality)

    def nieminen_centrality(self) -> np.ndarray:
        return self._vec(lib().hb_nieminen_centrality)

    def reachable_nodes(self) -> np.ndarray:
        return self._vec(lib().hb_reachable_nodes)

    def reachable_nodes_from(self, node: int) -> float:
        out = ctypes.c_double()
        _check(lib().hb_reachable_nodes_from(self._h, node, ctypes.byref(out)))
        return out.value

    # --- inspection ---
    def params(self) -> dict:
        p = _Params()
        _check(lib().hb_get_params(self._h, ctypes.byref(p)))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def counters(self, which: int = 0) -> np.ndarray:
        """Counter array in the reference layout, shape (n, words_per_counter)."""
        W = self.params()["words_per_counter"]
        out = np.zeros((self._n, W), dtype=np.uint64)
        _check(lib().hb_get_counters(self._h, which, _ptr(out), out.size))
        return out

    def counters_noalloc(self, which: int = 0) -> np.ndarray:
        """counters() through hb_get_counters_noalloc, which stages through the
        library's per-GPU array instead of allocating device memory."""
        W = self.params()["words_per_counter"]
        out = np.zeros((self._n, W), dtype=np.uint64)
        _check(lib().hb_get_counters_noalloc(self._h, which, _ptr(out), out.size))
        return out

    def modified_bits(self) -> np.ndarray:
        out = np.zeros((self._n + 63) // 64, dtype=np.uint64)
        _check(lib().hb_get_modified(self._h, _ptr(out), len(out)))
        return out

    def iteration_stats(self) -> list[dict]:
        k = ctypes.c_uint64()
        _check(lib().hb_num_iterations(self._h, ctypes.byref(k)))
        res = []
        for i in range(k.value):
            s = IterationStats()
            _check(lib().hb_get_iteration_stats(self._h, i, ctypes.byref(s)))
            res.append(s.as_dict())
        return res

    def set_timing(self, on: bool) -> None:
        _check(lib().hb_set_timing(self._h, int(on)))

    def autotune(self, reps: int = 2) -> tuple[int, list[tuple[int, float]]]:
        """Chooses the light/heavy row split by timing the first iteration
        that gathers whole counters (iteration 1, delta gathering off);
        returns (chosen threshold, [(threshold, ms), ...]).
        2**32 - 1 stands for "no heavy rows".  Resets the run state."""
        cap = 64
        th = np.zeros(cap, dtype=np.uint32)
        ms = np.zeros(cap)
        ln = ctypes.c_size_t()
        chosen = ctypes.c_uint32()
        _check(lib().hb_autotune(self._h, reps, _ptr(th), _ptr(ms), cap, ctypes.byref(ln),
                                 ctypes.byref(chosen)))
        k = min(ln.value, cap)
        return chosen.value, [(int(th[i]), float(ms[i])) for i in range(k)]

    def set_progress(self, f: Optional[Callable[[dict], None]]) -> None:
        """f(stats dict) after every iteration (None: off)."""
        if f is None:
            self._progress = None
            _check(lib().hb_set_progress(self._h, None, None))
            return

        def cb(stats_ptr, _ctx):
            f(stats_ptr.contents.as_dict())
        self._progress = PROGRESS_FN(cb)
        _check(lib().hb_set_progress(self._h, ctypes.cast(self._progress, ctypes.c_void_p), None))

    def set_stream(self, stream: Optional[int]) -> None:
        """Run on the caller's HIP stream (an integer handle, such as
        torch.cuda.Stream().cuda_stream); None gives back a private stream."""
        _check(lib().hb_set_stream(self._h, ctypes.c_void_p(stream) if stream else None))

    def set_delta_gathering(self, mode: int) -> None:
        """0: per iteration by the library's rule, 1: always, 2: never."""
        _check(lib().hb_set_delta_gathering(self._h, int(mode)))

    def set_window(self, mode: int, block_nodes: int = 0, half_width: int = 0,
                   heavy_degree: int = 0) -> None:
        """Window kernel for log2m-6 counters: 0 by the library's rule (the
        iterations that check every node and large worklists), 1 whenever it
        can run, 2 never; block_nodes / half_width cap its block and tile (0:
        the kernel's own); rows longer than heavy_degree take a workgroup (0:
        the default, 2**32 - 1: none) (hb_set_window)."""
        _check(lib().hb_set_window(self._h, int(mode), int(block_nodes), int(half_width),
                                   int(heavy_degree)))

    def set_streaming(self, mode: int) -> None:
        """Streamed counters: 0 by the per-iteration rule, 1 always streamed,
        2 always read in place over the host link (hb_set_streaming)."""
        _check(lib().hb_set_streaming(self._h, int(mode)))

    def set_tuning(self, heavy_degree: int) -> None:
        """Rows longer than heavy_degree go to the workgroup-per-node kernel
        (0 = default, 2**32 - 1 = never)."""
        _check(lib().hb_set_tuning(self._h, int(heavy_degree)))


class HyperBallBatch:
    """Many independent HyperBall instances advanced together (one fused
    launch sequence per batched iteration, replayed from a HIP graph)."""

    def __init__(self, builders: list):
        k = len(builders)
        self._builders = builders
        self._cfgs = (_Config * k)(*[b._config() for b in builders])
        self._gs = (_Csr * k)(*[b._graph.csr() for b in builders])
        self._ts = (_Csr * k)(*[b._transpose.csr() if b._transpose is not None else _Csr() for b in builders])
        h = ctypes.c_void_p()
        _check(lib().hb_batch_create(k, self._gs, self._ts, self._cfgs, ctypes.byref(h)))
        self._h = h
        self._fin = weakref.finalize(self, lib().hb_batch_destroy, h)
        self.instances = []
        for i, b in enumerate(builders):
            ih = ctypes.c_void_p()
            _check(lib().hb_batch_instance(h, i, ctypes.byref(ih)))
            hb = HyperBall(ih, b._graph.num_nodes, owner=self)
            for f in b._discounts:
                hb._add_discount(f)
            self.instances.append(hb)

    def close(self) -> None:
        self.instances = []
        self._fin()

    def run(self, upper_bound: int, threshold: Optional[float] = None) -> None:
        _check(lib().hb_batch_run(self._h, min(int(upper_bound), 2**64 - 1),
                                  math.nan if threshold is None else float(threshold)))

    def run_until_done(self) -> None:
        _check(lib().hb_batch_run_until_done(self._h))

    def set_graph(self, on: bool) -> None:
        _check(lib().hb_batch_set_graph(self._h, int(on)))

    def autotune(self, reps: int = 2) -> tuple[int, list[tuple[int, float]]]:
        """One light/heavy split for all instances, timed on the second
        batched iteration (delta gathering off); returns (chosen,
        [(threshold, ms), ...])."""
        cap = 64
        th = np.zeros(cap, dtype=np.uint32)
        ms = np.zeros(cap)
        ln = ctypes.c_size_t()
        chosen = ctypes.c_uint32()
        _check(lib().hb_batch_autotune(self._h, reps, _ptr(th), _ptr(ms), cap, ctypes.byref(ln),
                                       ctypes.byref(chosen)))
        k = min(ln.value, cap)
        return chosen.value, [(int(th[i]), float(ms[i])) for i in range(k)]

    def step_times(self) -> np.ndarray:
        ln = ctypes.c_size_t()
        _check(lib().hb_batch_step_times(self._h, None, 0, ctypes.byref(ln)))
        out = np.zeros(ln.value)
        _check(lib().hb_batch_step_times(self._h, _ptr(out), len(out), ctypes.byref(ln)))
        return out
