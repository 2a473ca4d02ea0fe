"""ctypes binding of libbrainmatch.so (the C ABI in include/brainmatch.h).

Host-side mirror of the reference tools' operations for Python callers
(tests, bench): load graphs, set a matching, score it, evaluate swap and
3-vertex permutation deltas, accept disjoint swaps, run the search drivers.
numpy arrays in, numpy arrays out; every non-zero status raises BMError with
the library's message.  There is no fallback: if the shared library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbrainmatch.so")
BIN_DIR = os.path.join(_HERE, "bin")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C brain-match_amd")

_lib = C.CDLL(LIB_PATH)

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)

KERNEL_SWAP, KERNEL_ROT3, KERNEL_SCORE = 0, 1, 2
KERNEL_GAIN_BUILD, KERNEL_GAIN_SWAP, KERNEL_GAIN_ROT3, KERNEL_GAIN_UPDATE, KERNEL_AUX = 3, 4, 5, 6, 7
WEIGHTS_GEOMETRIC, WEIGHTS_LOGNORMAL = 1, 2
SEARCH_SWAP_WEIGHTED, SEARCH_SWAP_UNIFORM, SEARCH_ROT3 = 0, 1, 2


class BMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"brainmatch error {code}: {msg}")
        self.code = code


class _Edges(C.Structure):
    _fields_ = [("m", C.c_int64), ("src", i32p), ("dst", i32p), ("w", i32p)]


class _EdgeList(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int32), ("src", i32p), ("dst", i32p), ("w", i32p)]


class _SynthParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("mean_out_degree", C.c_double), ("alpha", C.c_double),
                ("weights", C.c_int32), ("modify_frac", C.c_double), ("shuffle_frac", C.c_double),
                ("seed", C.c_uint64)]


class _Stats(C.Structure):
    _fields_ = [("count", C.c_int64), ("sum", C.c_int64), ("max", C.c_int64), ("argmax", C.c_int64),
                ("n_pos", C.c_int64), ("n_zero", C.c_int64)]


class _Tuning(C.Structure):
    _fields_ = [("waves_per_block", C.c_int32), ("lds_slots", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("group_lanes", C.c_int32), ("unroll", C.c_int32)]


class _SearchOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("seed", C.c_uint64), ("time_limit_s", C.c_double),
                ("max_iters", C.c_int64), ("out_path", C.c_char_p), ("sync_path", C.c_char_p),
                ("sync_interval_s", C.c_double), ("verify_every", C.c_int64), ("verbose", C.c_int32),
                ("stop_flag", C.POINTER(C.c_int)), ("restart_gap", C.c_double)]


class _SearchResult(C.Structure):
    _fields_ = [("initial_score", C.c_int64), ("best_score", C.c_int64), ("final_score", C.c_int64),
                ("iterations", C.c_int64), ("accepted", C.c_int64), ("evaluations", C.c_int64),
                ("verify_failures", C.c_int64), ("seconds", C.c_double), ("fallback_rounds", C.c_int64),
                ("restarts", C.c_int64)]


class _SweepOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("initial_swaps", C.c_int32), ("restart_swaps", C.c_int32),
                ("batch", C.c_int32), ("verify", C.c_int32), ("max_passes", C.c_int64), ("time_limit_s", C.c_double)]


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_c = C.c_void_p
_sig("bm_api_version", C.c_int, [])
_sig("bm_last_error", C.c_char_p, [_c])
_sig("bm_load_graph_csv", C.c_int, [C.c_char_p, C.POINTER(_EdgeList)])
_sig("bm_write_graph_csv", C.c_int, [C.c_char_p, C.POINTER(_EdgeList)])
_sig("bm_edge_list_free", None, [C.POINTER(_EdgeList)])
_sig("bm_load_mapping_csv", C.c_int, [C.c_char_p, C.c_int32, i32p, i64p])
_sig("bm_write_mapping_csv", C.c_int, [C.c_char_p, C.c_int32, i32p])
_sig("bm_synth_pair", C.c_int, [C.POINTER(_SynthParams), C.POINTER(_EdgeList), C.POINTER(_EdgeList), i32p, i32p])
_sig("bm_greedy_match", C.c_int, [C.c_int32, C.POINTER(_Edges), C.POINTER(_Edges), i32p])
_sig("bm_create", C.c_int, [C.POINTER(_c), C.c_int32, C.POINTER(_Edges), C.POINTER(_Edges), C.c_int32])
_sig("bm_destroy", None, [_c])
_sig("bm_num_vertices", C.c_int32, [_c])
_sig("bm_num_edges", C.c_int64, [_c, C.c_int])
_sig("bm_degrees", C.c_int, [_c, C.c_int, i32p, i32p])
_sig("bm_set_mapping", C.c_int, [_c, i32p])
_sig("bm_get_mapping", C.c_int, [_c, i32p])
_sig("bm_score", C.c_int, [_c, i64p])
_sig("bm_swap_deltas", C.c_int, [_c, i32p, C.c_int64, i64p])
_sig("bm_swap_deltas_changed", C.c_int, [_c, i32p, C.c_int32, i64p, i32p, i64p])
_sig("bm_rot3_deltas", C.c_int, [_c, i32p, C.c_int64, i64p, u8p])
_sig("bm_rot3_deltas_changed", C.c_int, [_c, i32p, C.c_int32, i64p, i32p, i64p])
_sig("bm_accept_disjoint", C.c_int, [_c, i32p, i64p, C.c_int64, i32p, i64p, i64p])
_sig("bm_apply_swap", C.c_int, [_c, C.c_int32, C.c_int32])
_sig("bm_apply_perm3", C.c_int, [_c, C.c_int32, C.c_int32, C.c_int32, C.c_int32])
_sig("bm_batch_create", C.c_int, [_c, C.c_int64, C.POINTER(_c)])
_sig("bm_batch_destroy", None, [_c])
_sig("bm_batch_size", C.c_int64, [_c])
_sig("bm_batch_set_pairs", C.c_int, [_c, i32p, C.c_int64])
_sig("bm_batch_set_triples", C.c_int, [_c, i32p, C.c_int64])
_sig("bm_batch_gen_pairs", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_batch_gen_triples", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_batch_gen_changed_pairs", C.c_int, [_c, i32p, C.c_int32])
_sig("bm_batch_gen_changed_triples", C.c_int, [_c, i32p, C.c_int32])
_sig("bm_batch_eval_swaps", C.c_int, [_c])
_sig("bm_batch_eval_swaps_uniform", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_batch_eval_rot3", C.c_int, [_c])
_sig("bm_batch_gen_umajor", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_batch_eval_swaps_umajor", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_batch_get", C.c_int, [_c, C.c_int64, C.c_int64, i64p, u8p, i32p])
_sig("bm_batch_stats", C.c_int, [_c, C.POINTER(_Stats)])
_sig("bm_batch_pick", C.c_int, [_c, C.c_int32, C.c_int64, i64p])
_sig("bm_batch_accept_disjoint", C.c_int, [_c, C.c_int64, i64p, i64p])
_sig("bm_timing_enable", C.c_int, [_c, C.c_int])
_sig("bm_timing_get", C.c_int, [_c, C.c_int32, i64p, C.POINTER(C.c_double)])
_sig("bm_timing_reset", C.c_int, [_c])
_sig("bm_sync", C.c_int, [_c])
_sig("bm_batch_algo_bytes", C.c_int, [_c, C.c_int32, i64p, i64p])
_sig("bm_uniform_algo_bytes", C.c_int, [_c, C.c_uint64, C.c_int64, i64p, i64p])
_sig("bm_umajor_algo_bytes", C.c_int, [_c, C.c_uint64, C.c_int64, i64p, i64p])
_sig("bm_get_tuning", C.c_int, [_c, C.c_int32, C.POINTER(_Tuning)])
_sig("bm_set_tuning", C.c_int, [_c, C.c_int32, C.POINTER(_Tuning)])
_sig("bm_device_info", C.c_int, [_c, C.c_char_p, C.c_size_t])
_sig("bm_search", C.c_int, [_c, C.POINTER(_SearchOpts), i32p, C.POINTER(_SearchResult)])
_sig("bm_hillclimb", C.c_int, [_c, C.c_uint64, C.c_int64, C.c_int64, C.c_int32, i64p, C.POINTER(_SearchResult)])
_sig("bm_gain_enable", C.c_int, [_c, C.c_int])
_sig("bm_gain_enabled", C.c_int, [_c])
_sig("bm_gain_dense_rows", C.c_int, [_c])
_sig("bm_gain_rebuild", C.c_int, [_c])
_sig("bm_batch_gen_amajor", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_batch_eval_rot3_amajor", C.c_int, [_c, C.c_uint64, C.c_int64])
_sig("bm_gain_table_bytes", C.c_int64, [C.c_int32])
_sig("bm_gain_row", C.c_int, [_c, C.c_int32, i32p])
_sig("bm_best_swap", C.c_int, [_c, i64p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)])
_sig("bm_best_swap_climb", C.c_int, [_c, C.c_int64, C.c_int64, i64p, C.POINTER(_SearchResult)])
_sig("bm_best_partners", C.c_int, [_c, C.c_int32, C.c_int32, i64p, i32p])
_sig("bm_sweep", C.c_int, [_c, C.POINTER(_SweepOpts), i32p, i64p, C.POINTER(_SearchResult)])
_sig("bm_twohop_set", C.c_int, [_c, C.c_int32, i32p, C.POINTER(C.c_int32)])
_sig("bm_batch_gen_twohop", C.c_int, [_c, i32p, C.c_int32, C.c_int32, C.c_uint64])
_sig("bm_hillclimb_twohop", C.c_int, [_c, C.c_uint64, C.c_int64, C.c_int32, C.c_int32, C.c_int32, i64p,
                                      C.POINTER(_SearchResult)])


def _p32(a):
    return a.ctypes.data_as(i32p) if a is not None else None


def _p64(a):
    return a.ctypes.data_as(i64p) if a is not None else None


def _pu8(a):
    return a.ctypes.data_as(u8p) if a is not None else None


def _check(rc, ctx=None):
    if rc != 0:
        msg = _lib.bm_last_error(ctx)
        raise BMError(rc, msg.decode() if msg else "")


def api_version():
    return _lib.bm_api_version()


@dataclass
class Graph:
    """Edge list of one directed weighted graph (ids 1..n)."""
    src: np.ndarray
    dst: np.ndarray
    w: np.ndarray
    n: int

    @property
    def m(self):
        return int(self.src.shape[0])

    def _edges(self):
        self.src = np.ascontiguousarray(self.src, dtype=np.int32)
        self.dst = np.ascontiguousarray(self.dst, dtype=np.int32)
        self.w = np.ascontiguousarray(self.w, dtype=np.int32)
        return _Edges(self.m, _p32(self.src), _p32(self.dst), _p32(self.w))


def _from_edge_list(el: _EdgeList) -> Graph:
    m = el.m
    src = np.ctypeslib.as_array(el.src, shape=(m,)).copy() if m else np.zeros(0, np.int32)
    dst = np.ctypeslib.as_array(el.dst, shape=(m,)).copy() if m else np.zeros(0, np.int32)
    w = np.ctypeslib.as_array(el.w, shape=(m,)).copy() if m else np.zeros(0, np.int32)
    g = Graph(src, dst, w, el.n)
    _lib.bm_edge_list_free(C.byref(el))
    return g


def load_graph_csv(path) -> Graph:
    el = _EdgeList()
    _check(_lib.bm_load_graph_csv(os.fsencode(path), C.byref(el)))
    return _from_edge_list(el)


def write_graph_csv(path, g: Graph):
    e = g._edges()
    el = _EdgeList(e.m, g.n, e.src, e.dst, e.w)
    _check(_lib.bm_write_graph_csv(os.fsencode(path), C.byref(el)))


def load_mapping_csv(path, n) -> np.ndarray:
    pi = np.zeros(n + 1, np.int32)
    lines = C.c_int64(0)
    _check(_lib.bm_load_mapping_csv(os.fsencode(path), n, _p32(pi), C.byref(lines)))
    return pi


def write_mapping_csv(path, pi):
    pi = np.ascontiguousarray(pi, dtype=np.int32)
    _check(_lib.bm_write_mapping_csv(os.fsencode(path), pi.shape[0] - 1, _p32(pi)))


def synth_pair(n, mean_out_degree, weights=WEIGHTS_GEOMETRIC, shuffle_frac=0.3, seed=1, alpha=2.1,
               modify_frac=0.2):
    """Returns (male, female, true_perm, pi0)."""
    p = _SynthParams(n, mean_out_degree, alpha, weights, modify_frac, shuffle_frac, seed)
    m, f = _EdgeList(), _EdgeList()
    truth = np.zeros(n + 1, np.int32)
    pi0 = np.zeros(n + 1, np.int32)
    _check(_lib.bm_synth_pair(C.byref(p), C.byref(m), C.byref(f), _p32(truth), _p32(pi0)))
    return _from_edge_list(m), _from_edge_list(f), truth, pi0


def greedy_match(n, male: Graph, female: Graph) -> np.ndarray:
    pi = np.zeros(n + 1, np.int32)
    em, ef = male._edges(), female._edges()
    _check(_lib.bm_greedy_match(n, C.byref(em), C.byref(ef), _p32(pi)))
    return pi


class Batch:
    def __init__(self, ctx: "Context", capacity: int):
        self.ctx = ctx
        h = C.c_void_p()
        _check(_lib.bm_batch_create(ctx.h, capacity, C.byref(h)), ctx.h)
        self.h = h
        self.capacity = capacity

    def close(self):
        if self.h:
            _lib.bm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        _check(rc, self.ctx.h)

    @property
    def size(self):
        return _lib.bm_batch_size(self.h)

    def set_pairs(self, uv):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1, 2)
        self._chk(_lib.bm_batch_set_pairs(self.h, _p32(uv), uv.shape[0]))

    def set_triples(self, abc):
        abc = np.ascontiguousarray(abc, dtype=np.int32).reshape(-1, 3)
        self._chk(_lib.bm_batch_set_triples(self.h, _p32(abc), abc.shape[0]))

    def gen_pairs(self, seed, k):
        self._chk(_lib.bm_batch_gen_pairs(self.h, seed, k))

    def gen_triples(self, seed, k):
        self._chk(_lib.bm_batch_gen_triples(self.h, seed, k))

    def gen_changed_pairs(self, changed):
        ch = np.ascontiguousarray(changed, dtype=np.int32)
        self._chk(_lib.bm_batch_gen_changed_pairs(self.h, _p32(ch), ch.shape[0]))

    def gen_changed_triples(self, changed):
        ch = np.ascontiguousarray(changed, dtype=np.int32)
        self._chk(_lib.bm_batch_gen_changed_triples(self.h, _p32(ch), ch.shape[0]))

    def gen_twohop(self, per_anchor, seed, anchors=None, n_anchors=None):
        """2-hop proposal pairs: per_anchor slots per anchor (explicit anchors,
        or n_anchors drawn from the seed)."""
        if anchors is not None:
            a = np.ascontiguousarray(anchors, dtype=np.int32)
            self._chk(_lib.bm_batch_gen_twohop(self.h, _p32(a), a.shape[0], per_anchor, seed))
        else:
            self._chk(_lib.bm_batch_gen_twohop(self.h, None, n_anchors, per_anchor, seed))

    def eval_swaps(self):
        self._chk(_lib.bm_batch_eval_swaps(self.h))

    def eval_swaps_uniform(self, seed, k):
        self._chk(_lib.bm_batch_eval_swaps_uniform(self.h, seed, k))

    def gen_umajor(self, seed, k):
        self._chk(_lib.bm_batch_gen_umajor(self.h, seed, k))

    def eval_swaps_umajor(self, seed, k):
        self._chk(_lib.bm_batch_eval_swaps_umajor(self.h, seed, k))

    def gen_amajor(self, seed, k):
        self._chk(_lib.bm_batch_gen_amajor(self.h, seed, k))

    def eval_rot3_amajor(self, seed, k):
        self._chk(_lib.bm_batch_eval_rot3_amajor(self.h, seed, k))

    def eval_rot3(self):
        self._chk(_lib.bm_batch_eval_rot3(self.h))

    def deltas(self, first=0, count=None):
        count = self.size - first if count is None else count
        d = np.empty(count, np.int64)
        self._chk(_lib.bm_batch_get(self.h, first, count, _p64(d), None, None))
        return d

    def perms(self, first=0, count=None):
        count = self.size - first if count is None else count
        p = np.empty(count, np.uint8)
        self._chk(_lib.bm_batch_get(self.h, first, count, None, _pu8(p), None))
        return p

    def pairs(self, first=0, count=None):
        count = self.size - first if count is None else count
        c = np.empty((count, 2), np.int32)
        self._chk(_lib.bm_batch_get(self.h, first, count, None, None, _p32(c)))
        return c

    def triples(self, first=0, count=None):
        count = self.size - first if count is None else count
        c = np.empty((count, 4), np.int32)
        self._chk(_lib.bm_batch_get(self.h, first, count, None, None, _p32(c)))
        return c

    def stats(self):
        s = _Stats()
        self._chk(_lib.bm_batch_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def pick(self, kind, rank):
        idx = C.c_int64(-1)
        self._chk(_lib.bm_batch_pick(self.h, kind, rank, C.byref(idx)))
        return idx.value

    def accept_disjoint(self, max_accept=0):
        na, gain = C.c_int64(0), C.c_int64(0)
        self._chk(_lib.bm_batch_accept_disjoint(self.h, max_accept, C.byref(na), C.byref(gain)))
        return na.value, gain.value

    def algo_bytes(self, generated):
        b, d = C.c_int64(0), C.c_int64(0)
        self._chk(_lib.bm_batch_algo_bytes(self.h, 1 if generated else 0, C.byref(b), C.byref(d)))
        return b.value, d.value


class Context:
    """Graphs + matching resident on one device (bm_create)."""

    def __init__(self, n, male: Graph, female: Graph, device=0):
        self.n = int(n)
        h = C.c_void_p()
        em, ef = male._edges(), female._edges()
        _check(_lib.bm_create(C.byref(h), self.n, C.byref(em), C.byref(ef), device))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.bm_destroy(self.h)
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

    def _chk(self, rc):
        _check(rc, self.h)

    def num_edges(self, female=False):
        return _lib.bm_num_edges(self.h, 1 if female else 0)

    def degrees(self, female=False):
        o = np.zeros(self.n + 1, np.int32)
        i = np.zeros(self.n + 1, np.int32)
        self._chk(_lib.bm_degrees(self.h, 1 if female else 0, _p32(o), _p32(i)))
        return o, i

    def set_mapping(self, pi):
        pi = np.ascontiguousarray(pi, dtype=np.int32)
        if pi.shape[0] != self.n + 1:
            raise ValueError(f"mapping must have n+1={self.n + 1} entries")
        self._chk(_lib.bm_set_mapping(self.h, _p32(pi)))

    def get_mapping(self):
        pi = np.zeros(self.n + 1, np.int32)
        self._chk(_lib.bm_get_mapping(self.h, _p32(pi)))
        return pi

    def score(self):
        s = C.c_int64(0)
        self._chk(_lib.bm_score(self.h, C.byref(s)))
        return s.value

    def swap_deltas(self, uv):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1, 2)
        d = np.empty(uv.shape[0], np.int64)
        self._chk(_lib.bm_swap_deltas(self.h, _p32(uv), uv.shape[0], _p64(d)))
        return d

    def swap_deltas_changed(self, changed):
        ch = np.ascontiguousarray(changed, dtype=np.int32)
        c = ch.shape[0]
        P = c * (c - 1) // 2 + c * (self.n - c)
        d = np.empty(P, np.int64)
        uv = np.empty((P, 2), np.int32)
        cnt = C.c_int64(0)
        self._chk(_lib.bm_swap_deltas_changed(self.h, _p32(ch), c, _p64(d), _p32(uv), C.byref(cnt)))
        return d, uv

    def rot3_deltas(self, abc):
        abc = np.ascontiguousarray(abc, dtype=np.int32).reshape(-1, 3)
        b = np.empty(abc.shape[0], np.int64)
        p = np.empty(abc.shape[0], np.uint8)
        self._chk(_lib.bm_rot3_deltas(self.h, _p32(abc), abc.shape[0], _p64(b), _pu8(p)))
        return b, p

    def rot3_deltas_changed(self, changed):
        ch = np.ascontiguousarray(changed, dtype=np.int32)
        c, n = ch.shape[0], self.n
        T = c * (c - 1) * (c - 2) // 6 + c * (c - 1) // 2 * (n - c) + (c * (n - c - 1) if n - c - 1 > 0 else 0)
        b = np.empty(T, np.int64)
        q = np.empty((T, 4), np.int32)
        cnt = C.c_int64(0)
        self._chk(_lib.bm_rot3_deltas_changed(self.h, _p32(ch), c, _p64(b), _p32(q), C.byref(cnt)))
        return b, q

    def accept_disjoint(self, uv, d):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1, 2)
        d = np.ascontiguousarray(d, dtype=np.int64)
        acc = np.empty(max(1, uv.shape[0]), np.int32)
        na, gain = C.c_int64(0), C.c_int64(0)
        self._chk(_lib.bm_accept_disjoint(self.h, _p32(uv), _p64(d), uv.shape[0], _p32(acc), C.byref(na),
                                          C.byref(gain)))
        return acc[:na.value].copy(), gain.value

    def apply_swap(self, u, v):
        self._chk(_lib.bm_apply_swap(self.h, u, v))

    def apply_perm3(self, a, b, c, p):
        self._chk(_lib.bm_apply_perm3(self.h, a, b, c, p))

    def batch(self, capacity):
        return Batch(self, capacity)

    def timing_enable(self, on=True):
        self._chk(_lib.bm_timing_enable(self.h, 1 if on else 0))

    def timing(self, kernel):
        n, ms = C.c_int64(0), C.c_double(0)
        self._chk(_lib.bm_timing_get(self.h, kernel, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def timing_reset(self):
        self._chk(_lib.bm_timing_reset(self.h))

    def sync(self):
        self._chk(_lib.bm_sync(self.h))

    def uniform_algo_bytes(self, seed, k):
        b, d = C.c_int64(0), C.c_int64(0)
        self._chk(_lib.bm_uniform_algo_bytes(self.h, seed, k, C.byref(b), C.byref(d)))
        return b.value, d.value

    def umajor_algo_bytes(self, seed, k):
        b, d = C.c_int64(0), C.c_int64(0)
        self._chk(_lib.bm_umajor_algo_bytes(self.h, seed, k, C.byref(b), C.byref(d)))
        return b.value, d.value

    def get_tuning(self, kernel):
        t = _Tuning()
        self._chk(_lib.bm_get_tuning(self.h, kernel, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _Tuning._fields_}

    def set_tuning(self, kernel, waves_per_block, lds_slots, blocks_per_cu, group_lanes=64, unroll=8):
        t = _Tuning(waves_per_block, lds_slots, blocks_per_cu, group_lanes, unroll)
        self._chk(_lib.bm_set_tuning(self.h, kernel, C.byref(t)))

    def device_info(self):
        buf = C.create_string_buffer(256)
        self._chk(_lib.bm_device_info(self.h, buf, 256))
        return buf.value.decode()

    def search(self, mode, seed=1, time_limit_s=0.0, max_iters=0, out_path=None, verify_every=64, verbose=0,
               restart_gap=0.002, sync_path=None, sync_interval_s=0.0):
        o = _SearchOpts(mode, seed, time_limit_s, max_iters, os.fsencode(out_path) if out_path else None,
                        os.fsencode(sync_path) if sync_path else None, sync_interval_s, verify_every, verbose, None,
                        restart_gap)
        best = np.zeros(self.n + 1, np.int32)
        r = _SearchResult()
        self._chk(_lib.bm_search(self.h, C.byref(o), _p32(best), C.byref(r)))
        return best, {k: getattr(r, k) for k, _ in _SearchResult._fields_}

    def twohop_set(self, v):
        out = np.zeros(self.n, np.int32)
        cnt = C.c_int32(0)
        self._chk(_lib.bm_twohop_set(self.h, v, _p32(out), C.byref(cnt)))
        return out[:cnt.value].copy()

    def greedy_match(self):
        """The greedy initial matching computed on the device (the same as
        greedy_match on the context's edges)."""
        pi = np.zeros(self.n + 1, np.int32)
        self._chk(_lib.bm_greedy_match_device(self.h, _p32(pi)))
        return pi

    def gain_enable(self, on=True):
        """Switch evaluation to the placement-gain table (or back to edge scans)."""
        self._chk(_lib.bm_gain_enable(self.h, 1 if on else 0))

    def gain_enabled(self):
        return bool(_lib.bm_gain_enabled(self.h))

    def gain_dense_rows(self):
        """Female rows the table build reads as dense 16-bit column arrays."""
        return int(_lib.bm_gain_dense_rows(self.h))

    def gain_rebuild(self):
        self._chk(_lib.bm_gain_rebuild(self.h))

    def gain_row(self, u):
        row = np.zeros(self.n + 1, np.int32)
        self._chk(_lib.bm_gain_row(self.h, u, _p32(row)))
        return row

    def best_swap(self):
        """(delta, u, v) of the best swap over all pairs u < v (needs the table)."""
        d = np.zeros(1, np.int64)
        u, v = C.c_int32(0), C.c_int32(0)
        self._chk(_lib.bm_best_swap(self.h, _p64(d), C.byref(u), C.byref(v)))
        return int(d[0]), u.value, v.value

    def best_swap_climb(self, max_iters, verify_every=1):
        trace = np.zeros(max_iters + 1, np.int64)
        r = _SearchResult()
        self._chk(_lib.bm_best_swap_climb(self.h, max_iters, verify_every, _p64(trace), C.byref(r)))
        res = {k: getattr(r, k) for k, _ in _SearchResult._fields_}
        return trace[:res["iterations"] + 1], res

    def best_partners(self, u0, count):
        """(delta, v) of the best partner v > u of each u in u0 .. u0+count-1."""
        d = np.zeros(count, np.int64)
        v = np.zeros(count, np.int32)
        self._chk(_lib.bm_best_partners(self.h, u0, count, _p64(d), _p32(v)))
        return d, v

    def sweep(self, seed=0, max_passes=0, time_limit_s=0.0, initial_swaps=50, restart_swaps=200, batch=0,
              verify=True):
        """The per-vertex best-partner sweep with restarts; returns (best
        matching, best score after each pass, result)."""
        o = _SweepOpts(seed, initial_swaps, restart_swaps, batch, 1 if verify else 0, max_passes, time_limit_s)
        best = np.zeros(self.n + 1, np.int32)
        trace = np.zeros(max(max_passes, 1), np.int64)
        r = _SearchResult()
        self._chk(_lib.bm_sweep(self.h, C.byref(o), _p32(best), _p64(trace) if max_passes > 0 else None,
                                C.byref(r)))
        res = {k: getattr(r, k) for k, _ in _SearchResult._fields_}
        return best, trace[:res["iterations"]] if max_passes > 0 else trace[:0], res

    def hillclimb_twohop(self, seed, rounds, anchors, per_anchor, verify=True):
        trace = np.zeros(rounds + 1, np.int64)
        r = _SearchResult()
        self._chk(_lib.bm_hillclimb_twohop(self.h, seed, rounds, anchors, per_anchor, 1 if verify else 0,
                                           _p64(trace), C.byref(r)))
        return trace, {k: getattr(r, k) for k, _ in _SearchResult._fields_}

    def hillclimb(self, seed, rounds, batch, verify=True):
        trace = np.zeros(rounds + 1, np.int64)
        r = _SearchResult()
        self._chk(_lib.bm_hillclimb(self.h, seed, rounds, batch, 1 if verify else 0, _p64(trace), C.byref(r)))
        return trace, {k: getattr(r, k) for k, _ in _SearchResult._fields_}


class _AnnealParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("chains_per_instance", C.c_int32), ("t0", C.c_double),
                ("decay", C.c_double), ("p_rot3", C.c_double), ("traj_cap", C.c_int32), ("p_pool", C.c_double)]


_sig("bm_greedy_match_device", C.c_int, [_c, i32p])
_sig("bm_anneal_create", C.c_int, [C.POINTER(_c), C.POINTER(_c), C.c_int32, C.POINTER(_AnnealParams)])
_sig("bm_anneal_destroy", None, [_c])
_sig("bm_anneal_num_chains", C.c_int64, [_c])
_sig("bm_anneal_run", C.c_int, [_c, C.c_int64, C.c_int32])
_sig("bm_anneal_state", C.c_int, [_c, i64p, i64p, i64p, i64p])
_sig("bm_anneal_mapping", C.c_int, [_c, C.c_int64, C.c_int32, i32p])
_sig("bm_anneal_trajectory", C.c_int, [_c, i64p, C.POINTER(C.c_int32)])
_sig("bm_anneal_verify", C.c_int, [_c, i64p])
_sig("bm_anneal_timing", C.c_int, [_c, i64p, C.POINTER(C.c_double)])
_sig("bm_anneal_algo_bytes", C.c_int, [_c, i64p, i64p])
_sig("bm_anneal_last_error", C.c_char_p, [_c])
_sig("bm_anneal_schedule", C.c_int, [_c, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)])
_sig("bm_anneal_reset", C.c_int, [_c])
_sig("bm_anneal_move_counts", C.c_int, [_c, i64p, i64p])
_sig("bm_anneal_set_pool", C.c_int, [_c, C.c_int32, i32p, C.c_int64])
_sig("bm_anneal_pool_learned", C.c_int, [_c, C.c_int32, _c, C.c_uint64, C.c_int64, C.c_int64])
_sig("bm_anneal_get_pool", C.c_int, [_c, C.c_int32, i32p, i64p])
_sig("bm_anneal_p_pool", C.c_uint64, [_c])
_sig("bm_anneal_get_tuning", C.c_int, [_c, C.POINTER(_Tuning)])
_sig("bm_anneal_set_tuning", C.c_int, [_c, C.POINTER(_Tuning)])
_sig("bm_anneal_autotune", C.c_int, [_c, C.c_int64, C.POINTER(_Tuning), C.POINTER(C.c_double), C.c_int32,
                                     C.POINTER(C.c_int32)])


class Anneal:
    """Annealing chains over one or more contexts (instances) on one device:
    chains_per_instance chains per context, all advanced by one kernel launch
    replayed from a HIP graph (bm_anneal_* in include/brainmatch.h)."""

    def __init__(self, contexts, chains_per_instance=1, seed=1, t0=50.0, decay=0.9999, p_rot3=0.0, traj_cap=0,
                 p_pool=0.0):
        self.contexts = list(contexts)
        self.seed = seed
        self.ns = [c.n for c in self.contexts]
        self.per_instance = chains_per_instance
        self.traj_cap = traj_cap
        arr = (_c * len(self.contexts))(*[c.h for c in self.contexts])
        p = _AnnealParams(seed, chains_per_instance, t0, decay, p_rot3, traj_cap, p_pool)
        self.h = _c()
        rc = _lib.bm_anneal_create(C.byref(self.h), arr, len(self.contexts), C.byref(p))
        if rc != 0:
            msg = _lib.bm_anneal_last_error(None)
            raise BMError(rc, msg.decode() if msg else "")
        self.chains = _lib.bm_anneal_num_chains(self.h)

    def close(self):
        if getattr(self, "h", None):
            _lib.bm_anneal_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            msg = _lib.bm_anneal_last_error(self.h)
            raise BMError(rc, msg.decode() if msg else "")

    def run(self, steps, launches=1):
        self._chk(_lib.bm_anneal_run(self.h, steps, launches))

    def state(self):
        k = self.chains
        s, b, st, a = (np.zeros(k, np.int64) for _ in range(4))
        self._chk(_lib.bm_anneal_state(self.h, _p64(s), _p64(b), _p64(st), _p64(a)))
        return dict(score=s, best=b, steps=st, accepted=a)

    def instance_of(self, chain):
        return chain // self.per_instance

    def mapping(self, chain, best=False):
        inst = self.instance_of(chain) if 0 <= chain < self.chains else 0   # the library rejects bad chains
        pi = np.zeros(self.ns[inst] + 1, np.int32)
        self._chk(_lib.bm_anneal_mapping(self.h, chain, 1 if best else 0, _p32(pi)))
        return pi

    def trajectory(self):
        t = np.zeros((self.chains, max(1, self.traj_cap)), np.int64)
        cnt = C.c_int32(0)
        self._chk(_lib.bm_anneal_trajectory(self.h, _p64(t), C.byref(cnt)))
        return t[:, :cnt.value].copy()

    def verify(self):
        bad = C.c_int64(0)
        self._chk(_lib.bm_anneal_verify(self.h, C.byref(bad)))
        return bad.value

    def timing(self):
        n, ms = C.c_int64(0), C.c_double(0)
        self._chk(_lib.bm_anneal_timing(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def algo_bytes(self):
        b, s = C.c_int64(0), C.c_int64(0)
        self._chk(_lib.bm_anneal_algo_bytes(self.h, C.byref(b), C.byref(s)))
        return b.value, s.value

    def reset(self):
        self._chk(_lib.bm_anneal_reset(self.h))

    def move_counts(self):
        """Per chain: 3-vertex moves and moves built on a pool pair so far."""
        r = np.zeros(self.chains, np.int64)
        p = np.zeros(self.chains, np.int64)
        self._chk(_lib.bm_anneal_move_counts(self.h, _p64(r), _p64(p)))
        return r, p

    def set_pool(self, inst, uv):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1, 2)
        self._chk(_lib.bm_anneal_set_pool(self.h, inst, _p32(uv), uv.shape[0]))

    def pool_learned(self, inst, mlp, seed, pool=1 << 20, top=1 << 16):
        """Instance inst's pool <- the top best-ranked of `pool` uniform pairs
        under mlp (a network of that instance's context)."""
        self._chk(_lib.bm_anneal_pool_learned(self.h, inst, mlp.h, seed, pool, top))

    def get_pool(self, inst):
        k = C.c_int64(0)
        self._chk(_lib.bm_anneal_get_pool(self.h, inst, None, C.byref(k)))
        uv = np.zeros((k.value, 2), np.int32)
        self._chk(_lib.bm_anneal_get_pool(self.h, inst, _p32(uv), C.byref(k)))
        return uv

    def p_pool(self):
        """The pool threshold on 32 random bits, as the kernels use it."""
        return int(_lib.bm_anneal_p_pool(self.h))

    def tuning(self):
        t = _Tuning()
        self._chk(_lib.bm_anneal_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _Tuning._fields_}

    def set_tuning(self, waves_per_block, lds_slots, blocks_per_cu, unroll=4):
        t = _Tuning(waves_per_block, lds_slots, blocks_per_cu, 64, unroll)
        self._chk(_lib.bm_anneal_set_tuning(self.h, C.byref(t)))

    def autotune(self, steps=100):
        best = _Tuning()
        trials = np.zeros((96, 5), np.float64)   # waves, slots, blocks_per_cu, unroll, ms
        nt = C.c_int32(0)
        self._chk(_lib.bm_anneal_autotune(self.h, steps, C.byref(best), trials.ctypes.data_as(C.POINTER(C.c_double)),
                                          96, C.byref(nt)))
        return {k: getattr(best, k) for k, _ in _Tuning._fields_}, trials[:nt.value].copy()

    def schedule(self):
        b, l, p = C.c_double(0), C.c_double(0), C.c_uint64(0)
        self._chk(_lib.bm_anneal_schedule(self.h, C.byref(b), C.byref(l), C.byref(p)))
        return b.value, l.value, p.value


PROPOSAL_UNIFORM, PROPOSAL_TWOHOP, PROPOSAL_LEARNED = 0, 1, 2


class _ClimbOpts(C.Structure):
    _fields_ = [("proposal", C.c_int32), ("seed", C.c_uint64), ("time_limit_s", C.c_double),
                ("max_rounds", C.c_int64), ("per_round", C.c_int64), ("pool", C.c_int64),
                ("per_anchor", C.c_int32), ("mlp", _c), ("refresh_every", C.c_int32), ("verify", C.c_int32),
                ("retrain_every", C.c_int32), ("retrain_lr", C.c_float), ("fallback", C.c_int32)]


_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_sig("bm_mlp_num_params", C.c_int32, [])
_sig("bm_mlp_create", C.c_int, [_c, C.c_uint64, C.POINTER(_c)])
_sig("bm_mlp_destroy", None, [_c])
_sig("bm_mlp_get_params", C.c_int, [_c, _fp])
_sig("bm_mlp_set_params", C.c_int, [_c, _fp])
_sig("bm_mlp_refresh_features", C.c_int, [_c])
_sig("bm_mlp_vertex_features", C.c_int, [_c, _fp])
_sig("bm_mlp_score_pairs", C.c_int, [_c, i32p, C.c_int64, _fp])
_sig("bm_mlp_gradient", C.c_int, [_c, i32p, _fp, C.c_int64, _fp, _fp])
_sig("bm_mlp_train", C.c_int, [_c, C.c_uint64, C.c_int64, C.c_int32, C.c_float, _fp])
_sig("bm_mlp_train_batch", C.c_int, [_c, _c, C.c_float, _fp])
_sig("bm_batch_gen_learned", C.c_int, [_c, _c, C.c_uint64, C.c_int64, C.c_int64])
_sig("bm_batch_gen_learned_noalloc", C.c_int, [_c, _c, C.c_uint64, C.c_int64, C.c_int64])
_sig("bm_climb", C.c_int, [_c, C.POINTER(_ClimbOpts), i64p, _dp, C.c_int64, C.POINTER(_SearchResult)])


def _pf(a):
    return a.ctypes.data_as(_fp) if a is not None else None


class Mlp:
    """The learned swap proposal of a context (bm_mlp_* in include/brainmatch.h)."""

    N_PARAMS = int(_lib.bm_mlp_num_params())

    def __init__(self, ctx: "Context", seed=1):
        self.ctx = ctx
        self.h = _c()
        ctx._chk(_lib.bm_mlp_create(ctx.h, seed, C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None):
            _lib.bm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        self.ctx._chk(rc)

    def params(self):
        p = np.zeros(self.N_PARAMS, np.float32)
        self._chk(_lib.bm_mlp_get_params(self.h, _pf(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.N_PARAMS,)
        self._chk(_lib.bm_mlp_set_params(self.h, _pf(p)))

    @staticmethod
    def unpack(p):
        """(W1 16x128, b1, W2 128x128, b2, w3, b3) views of a parameter vector."""
        o = 0
        out = []
        for shape in ((16, 128), (128,), (128, 128), (128,), (128,), (1,)):
            size = int(np.prod(shape))
            out.append(p[o:o + size].reshape(shape))
            o += size
        return out

    def refresh_features(self):
        self._chk(_lib.bm_mlp_refresh_features(self.h))

    def vertex_features(self):
        out = np.zeros((self.ctx.n + 1, 8), np.float32)
        self._chk(_lib.bm_mlp_vertex_features(self.h, _pf(out)))
        return out

    def score_pairs(self, uv):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1, 2)
        out = np.zeros(uv.shape[0], np.float32)
        self._chk(_lib.bm_mlp_score_pairs(self.h, _p32(uv), uv.shape[0], _pf(out)))
        return out

    def gradient(self, uv, target):
        uv = np.ascontiguousarray(uv, dtype=np.int32).reshape(-1, 2)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.N_PARAMS, np.float32)
        mse = C.c_float(0)
        self._chk(_lib.bm_mlp_gradient(self.h, _p32(uv), _pf(t), uv.shape[0], _pf(g), C.byref(mse)))
        return g, mse.value

    def train(self, seed, steps, batch=1 << 14, lr=1e-3):
        tr = np.zeros(steps, np.float32)
        self._chk(_lib.bm_mlp_train(self.h, seed, steps, batch, lr, _pf(tr)))
        return tr

    def train_batch(self, batch: "Batch", lr=1e-3):
        """One Adam step on an evaluated pair batch's pairs and exact deltas;
        returns the batch loss before the step."""
        mse = C.c_float(0)
        self._chk(_lib.bm_mlp_train_batch(self.h, batch.h, lr, C.byref(mse)))
        return mse.value


def _batch_gen_learned(self, mlp, seed, pool, top):
    self._chk(_lib.bm_batch_gen_learned(self.h, mlp.h, seed, pool, top))


def _batch_gen_learned_noalloc(self, mlp, seed, pool, top):
    """gen_learned with no device allocation (fixed library scratch)."""
    self._chk(_lib.bm_batch_gen_learned_noalloc(self.h, mlp.h, seed, pool, top))


Batch.gen_learned = _batch_gen_learned
Batch.gen_learned_noalloc = _batch_gen_learned_noalloc


def _ctx_climb(self, proposal, seed=1, time_limit_s=0.0, max_rounds=0, per_round=4096, pool=1 << 20,
               per_anchor=64, mlp=None, refresh_every=1, verify=False, trace_cap=100000, retrain_every=0,
               retrain_lr=1e-3, fallback=False):
    o = _ClimbOpts(proposal, seed, time_limit_s, max_rounds, per_round, pool, per_anchor,
                   mlp.h if mlp is not None else None, refresh_every, 1 if verify else 0, retrain_every, retrain_lr,
                   1 if fallback else 0)
    ts = np.zeros(trace_cap, np.int64)
    tt = np.zeros(trace_cap, np.float64)
    r = _SearchResult()
    self._chk(_lib.bm_climb(self.h, C.byref(o), _p64(ts), tt.ctypes.data_as(_dp), trace_cap, C.byref(r)))
    res = {k: getattr(r, k) for k, _ in _SearchResult._fields_}
    n = min(trace_cap, res["iterations"])
    return ts[:n].copy(), tt[:n].copy(), res


Context.climb = _ctx_climb
