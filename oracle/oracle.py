"""ctypes binding of the CPU oracle (liboracle.so) and of the reference code
compiled into oracle/_ref.  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker, never
by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libref_gao.so")
REF_SCORE = os.path.join(REF_DIR, "score-matching")
REF_GREEDY = os.path.join(REF_DIR, "greedy-match")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run make -C oracle")
_o = C.CDLL(LIB_PATH)

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
vp = C.c_void_p


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


_sig(_o, "or_build", vp, [C.c_int32, C.c_int64, i32p, i32p, i32p, C.c_int64, i32p, i32p, i32p])
_sig(_o, "or_free", None, [vp])
_sig(_o, "or_num_edges", C.c_int64, [vp, C.c_int])
_sig(_o, "or_score", C.c_int64, [vp, i32p])
_sig(_o, "or_swap_delta", C.c_int64, [vp, i32p, C.c_int32, C.c_int32])
_sig(_o, "or_swap_deltas", None, [vp, i32p, i32p, C.c_int64, i64p])
_sig(_o, "or_swap_delta_brute", C.c_int64, [vp, i32p, C.c_int32, C.c_int32])
_sig(_o, "or_rot3_all", None, [vp, i32p, C.c_int32, C.c_int32, C.c_int32, C.c_int, i64p])
_sig(_o, "or_rot3", None, [vp, i32p, C.c_int32, C.c_int32, C.c_int32, i64p, i32p])
_sig(_o, "or_rot3_many", None, [vp, i32p, i32p, C.c_int64, i64p, i32p])
_sig(_o, "or_quick_enum", C.c_int64, [C.c_int32, i32p, C.c_int32, i32p])
_sig(_o, "or_three_enum", C.c_int64, [C.c_int32, i32p, C.c_int32, i32p])
_sig(_o, "or_dense_build", vp, [C.c_int32, C.c_int64, i32p, i32p, i32p, C.c_int64, i32p, i32p, i32p])
_sig(_o, "or_dense_free", None, [vp])
_sig(_o, "or_ref_dense_swap", C.c_int32, [vp, i32p, C.c_int32, C.c_int32])
_sig(_o, "or_ref_dense_quick", C.c_int64, [vp, i32p, i32p, C.c_int32, i32p, i32p])
_sig(_o, "or_ref_dense_three", None, [vp, i32p, C.c_int32, C.c_int32, C.c_int32, i32p, i32p])
_sig(_o, "or_ref_dense_three_launch", C.c_int64, [vp, i32p, i32p, C.c_int32, i32p, i32p])
_sig(_o, "or_lists_build", vp, [C.c_int32, C.c_int64, i32p, i32p, i32p, C.c_int64, i32p, i32p, i32p])
_sig(_o, "or_lists_free", None, [vp])
_sig(_o, "or_ref_lists_delta", C.c_int32, [vp, i32p, C.c_int32, C.c_int32])
_sig(_o, "or_ref_lists_deltas", None, [vp, i32p, i32p, C.c_int64, i32p, C.c_int])
_sig(_o, "or_max_threads", C.c_int, [])
_sig(_o, "or_uniform_pairs", None, [C.c_uint64, C.c_int64, C.c_int32, i32p])
_sig(_o, "or_accept_disjoint", C.c_int64, [vp, i32p, i32p, i64p, C.c_int64, C.c_int64, i64p, i64p])
_sig(_o, "or_hillclimb", None, [vp, i32p, C.c_uint64, C.c_int64, C.c_int64, i64p])
_sig(_o, "or_best_swap", None, [vp, i32p, i64p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)])
_sig(_o, "or_best_swap_climb", C.c_int64, [vp, i32p, C.c_int64, i64p])


def _a32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _p32(a):
    return a.ctypes.data_as(i32p)


def _p64(a):
    return a.ctypes.data_as(i64p)


def _edges(g):
    s, d, w = _a32(g.src), _a32(g.dst), _a32(g.w)
    return s, d, w


class Oracle:
    """Exact score / deltas on the CPU (score-matching semantics)."""

    def __init__(self, n, male, female):
        self.n = int(n)
        self._keep = _edges(male) + _edges(female)
        ms, md, mw, fs, fd, fw = self._keep
        self.h = _o.or_build(self.n, ms.shape[0], _p32(ms), _p32(md), _p32(mw), fs.shape[0], _p32(fs), _p32(fd),
                             _p32(fw))
        if not self.h:
            raise ValueError("oracle: invalid graph")

    def __del__(self):
        if getattr(self, "h", None):
            _o.or_free(self.h)
            self.h = None

    def num_edges(self, female=False):
        return _o.or_num_edges(self.h, 1 if female else 0)

    def score(self, pi):
        pi = _a32(pi)
        return _o.or_score(self.h, _p32(pi))

    def swap_delta(self, pi, u, v):
        pi = _a32(pi)
        return _o.or_swap_delta(self.h, _p32(pi), u, v)

    def swap_deltas(self, pi, uv):
        pi, uv = _a32(pi), _a32(uv).reshape(-1, 2)
        out = np.empty(uv.shape[0], np.int64)
        _o.or_swap_deltas(self.h, _p32(pi), _p32(uv), uv.shape[0], _p64(out))
        return out

    def swap_delta_brute(self, pi, u, v):
        pi = _a32(pi)
        return _o.or_swap_delta_brute(self.h, _p32(pi), u, v)

    def rot3_all(self, pi, a, b, c, brute=False):
        pi = _a32(pi)
        d5 = np.empty(5, np.int64)
        _o.or_rot3_all(self.h, _p32(pi), a, b, c, 1 if brute else 0, _p64(d5))
        return d5

    def rot3_many(self, pi, abc):
        pi, abc = _a32(pi), _a32(abc).reshape(-1, 3)
        best = np.empty(abc.shape[0], np.int64)
        perm = np.empty(abc.shape[0], np.int32)
        _o.or_rot3_many(self.h, _p32(pi), _p32(abc), abc.shape[0], _p64(best), _p32(perm))
        return best, perm

    def accept_disjoint(self, pi, uv, d, max_accept=0):
        pi = _a32(pi).copy()
        uv = _a32(uv).reshape(-1, 2)
        d = np.ascontiguousarray(d, dtype=np.int64)
        acc = np.empty(max(1, uv.shape[0]), np.int64)
        gain = C.c_int64(0)
        na = _o.or_accept_disjoint(self.h, _p32(pi), _p32(uv), _p64(d), uv.shape[0], max_accept, _p64(acc),
                                   C.byref(gain))
        return pi, acc[:na].copy(), gain.value

    def best_swap(self, pi):
        pi = _a32(pi)
        d = np.zeros(1, np.int64)
        u, v = C.c_int32(0), C.c_int32(0)
        _o.or_best_swap(self.h, _p32(pi), _p64(d), C.byref(u), C.byref(v))
        return int(d[0]), u.value, v.value

    def best_swap_climb(self, pi, max_iters):
        pi = _a32(pi).copy()
        trace = np.zeros(max_iters + 1, np.int64)
        k = _o.or_best_swap_climb(self.h, _p32(pi), max_iters, _p64(trace))
        return pi, trace[:k + 1]

    def hillclimb(self, pi, seed, rounds, batch):
        pi = _a32(pi).copy()
        trace = np.empty(rounds + 1, np.int64)
        _o.or_hillclimb(self.h, _p32(pi), seed, rounds, batch, _p64(trace))
        return pi, trace


def uniform_pairs(seed, k, n):
    uv = np.empty((k, 2), np.int32)
    _o.or_uniform_pairs(seed, k, n, _p32(uv))
    return uv


def quick_enum(n, changed):
    ch = _a32(changed)
    c = ch.shape[0]
    P = c * (c - 1) // 2 + c * (n - c)
    uv = np.empty((P, 2), np.int32)
    _o.or_quick_enum(n, _p32(ch), c, _p32(uv))
    return uv


def three_enum(n, changed):
    ch = _a32(changed)
    c = ch.shape[0]
    T = c * (c - 1) * (c - 2) // 6 + c * (c - 1) // 2 * (n - c) + (c * (n - c - 1) if n - c - 1 > 0 else 0)
    abc = np.empty((T, 3), np.int32)
    _o.or_three_enum(n, _p32(ch), c, _p32(abc))
    return abc


class DenseRef:
    """The reference kernels' arithmetic on dense int16 matrices."""

    def __init__(self, n, male, female):
        self.n = int(n)
        self._keep = _edges(male) + _edges(female)
        ms, md, mw, fs, fd, fw = self._keep
        self.h = _o.or_dense_build(self.n, ms.shape[0], _p32(ms), _p32(md), _p32(mw), fs.shape[0], _p32(fs),
                                   _p32(fd), _p32(fw))

    def __del__(self):
        if getattr(self, "h", None):
            _o.or_dense_free(self.h)
            self.h = None

    def swap(self, pi, u, v):
        pi = _a32(pi)
        return _o.or_ref_dense_swap(self.h, _p32(pi), u, v)

    def quick(self, pi, changed):
        pi, ch = _a32(pi), _a32(changed)
        c = ch.shape[0]
        P = c * (c - 1) // 2 + c * (self.n - c)
        d = np.empty(P, np.int32)
        uv = np.empty((P, 2), np.int32)
        _o.or_ref_dense_quick(self.h, _p32(pi), _p32(ch), c, _p32(d), _p32(uv))
        return d, uv

    def three(self, pi, a, b, c):
        pi = _a32(pi)
        best, p = C.c_int32(0), C.c_int32(0)
        _o.or_ref_dense_three(self.h, _p32(pi), a, b, c, C.byref(best), C.byref(p))
        return best.value, p.value

    def three_launch(self, pi, changed):
        pi, ch = _a32(pi), _a32(changed)
        c, n = ch.shape[0], self.n
        T = c * (c - 1) * (c - 2) // 6 + c * (c - 1) // 2 * (n - c) + (c * (n - c - 1) if n - c - 1 > 0 else 0)
        d = np.empty(T, np.int32)
        t4 = np.empty((T, 4), np.int32)
        _o.or_ref_dense_three_launch(self.h, _p32(pi), _p32(ch), c, _p32(d), _p32(t4))
        return d, t4


class ListsRef:
    """calculate_swap_delta's arithmetic (edge lists + dense female matrix)."""

    def __init__(self, n, male, female):
        self.n = int(n)
        self._keep = _edges(male) + _edges(female)
        ms, md, mw, fs, fd, fw = self._keep
        self.h = _o.or_lists_build(self.n, ms.shape[0], _p32(ms), _p32(md), _p32(mw), fs.shape[0], _p32(fs),
                                   _p32(fd), _p32(fw))
        if not self.h:
            raise MemoryError("oracle lists build failed")

    def __del__(self):
        if getattr(self, "h", None):
            _o.or_lists_free(self.h)
            self.h = None

    def delta(self, pi, u, v):
        pi = _a32(pi)
        return _o.or_ref_lists_delta(self.h, _p32(pi), u, v)

    def deltas(self, pi, uv, threads=0):
        pi, uv = _a32(pi), _a32(uv).reshape(-1, 2)
        out = np.empty(uv.shape[0], np.int32)
        _o.or_ref_lists_deltas(self.h, _p32(pi), _p32(uv), uv.shape[0], _p32(out), threads)
        return out


def max_threads():
    return _o.or_max_threads()


# ---------------------------------------------------------------------------
# oracle/_ref: the reference's own code, compiled from its sources here
# ---------------------------------------------------------------------------

def ref_available():
    return os.path.exists(REF_LIB) and os.path.exists(REF_SCORE)


class RefCode:
    """calculate_swap_delta / calculate_alignment_score of
    src/graph-alignment-optimized.c, compiled unmodified (ids <= 18524)."""

    def __init__(self, male, female):
        self.lib = C.CDLL(REF_LIB)
        _sig(self.lib, "refh_build", vp, [C.c_int64, i32p, i32p, i32p])
        _sig(self.lib, "refh_free", None, [vp])
        _sig(self.lib, "refh_score", C.c_int32, [vp, vp, i32p, C.c_int32])
        _sig(self.lib, "refh_deltas", None, [vp, vp, i32p, C.c_int32, i32p, C.c_int64, i32p, C.c_int])
        _sig(self.lib, "refh_num_nodes", C.c_int, [])
        self.cap = self.lib.refh_num_nodes()
        self._keep = _edges(male) + _edges(female)
        ms, md, mw, fs, fd, fw = self._keep
        self.gm = self.lib.refh_build(ms.shape[0], _p32(ms), _p32(md), _p32(mw))
        self.gf = self.lib.refh_build(fs.shape[0], _p32(fs), _p32(fd), _p32(fw))

    def close(self):
        if getattr(self, "gm", None):
            self.lib.refh_free(self.gm)
            self.lib.refh_free(self.gf)
            self.gm = self.gf = None

    def __del__(self):
        self.close()

    def score(self, pi):
        pi = _a32(pi)
        return self.lib.refh_score(self.gm, self.gf, _p32(pi), pi.shape[0] - 1)

    def deltas(self, pi, uv, threads=0):
        pi, uv = _a32(pi), _a32(uv).reshape(-1, 2)
        out = np.empty(uv.shape[0], np.int32)
        self.lib.refh_deltas(self.gm, self.gf, _p32(pi), pi.shape[0] - 1, _p32(uv), uv.shape[0], _p32(out), threads)
        return out


def ref_score_matching(g1_csv, g2_csv, matching_csv):
    """Runs the compiled score-matching and returns its integer score (the
    only stdout line that is a bare integer, src/score-matching.c:685)."""
    r = subprocess.run([REF_SCORE, g1_csv, g2_csv, matching_csv], capture_output=True, text=True, check=True)
    for line in r.stdout.splitlines():
        s = line.strip()
        if s.lstrip("-").isdigit():
            return int(s)
    raise RuntimeError("score-matching printed no score:\n" + r.stdout[-2000:])


_sig(_o, "or_gain_row", None, [vp, i32p, C.c_int32, i32p])


def gain_row(o: "Oracle", pi, u):
    """Row u of the placement-gain table from its definition (small graphs)."""
    pi = _a32(pi)
    row = np.empty(o.n + 1, np.int32)
    _o.or_gain_row(o.h, _p32(pi), u, _p32(row))
    return row


_sig(_o, "or_det_exp_many", None, [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double)])
_sig(_o, "or_uniform_triple", None, [C.c_uint64, C.c_uint64, C.c_int32, i32p, i32p, i32p])
_sig(_o, "or_anneal_chain", None, [vp, i32p, i32p, C.c_uint64, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                   C.c_uint64, i64p, i32p, C.c_int64, C.c_uint64])
_sig(_o, "or_chain_seed", C.c_uint64, [C.c_uint64, C.c_int64])


def det_exp(x):
    """The deterministic exp of the annealing acceptance test, elementwise."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    _o.or_det_exp_many(x.ctypes.data_as(C.POINTER(C.c_double)), x.size,
                       out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def uniform_triple(seed, c, n):
    a, b, d = C.c_int32(), C.c_int32(), C.c_int32()
    _o.or_uniform_triple(seed, c, n, C.byref(a), C.byref(b), C.byref(d))
    return a.value, b.value, d.value


_sig(_o, "or_best_partner", None, [vp, i32p, C.c_int32, i64p, i32p])
_sig(_o, "or_sweep_perturb", None, [i32p, C.c_int32, C.c_uint64, C.c_uint64, C.c_int32, i32p])
_sig(_o, "or_sweep", None, [vp, i32p, C.c_uint64, C.c_int64, C.c_int32, C.c_int32, i64p, i64p])


def best_partner(o: "Oracle", pi, u):
    """(delta, v) of the best partner v > u of u (ties to the smaller v)."""
    pi = _a32(pi)
    d, v = C.c_int64(), C.c_int32()
    _o.or_best_partner(o.h, _p32(pi), u, C.byref(d), C.byref(v))
    return d.value, v.value


def sweep_perturb(pi, n, seed, p, k):
    """The sweep's perturbation p of pi (a copy) and the 2k vertices it moved."""
    pi = _a32(pi).copy()
    sel = np.zeros(max(2 * k, 1), np.int32)
    _o.or_sweep_perturb(_p32(pi), n, seed, p, k, _p32(sel))
    return pi, sel[:2 * k]


def sweep(o: "Oracle", pi, seed, max_passes, initial_swaps=50, restart_swaps=200):
    """Replay of the per-vertex best-partner sweep: (best matching, best score
    after each pass, {best, accepted, restarts, evaluations})."""
    pi = _a32(pi).copy()
    trace = np.zeros(max_passes, np.int64)
    out = np.zeros(4, np.int64)
    _o.or_sweep(o.h, _p32(pi), seed, max_passes, initial_swaps, restart_swaps, _p64(trace), _p64(out))
    return pi, trace, dict(zip(("best_score", "accepted", "restarts", "evaluations"), (int(x) for x in out)))


def chain_seed(seed, index):
    return int(_o.or_chain_seed(seed, index))


_sig(_o, "or_third_vertex", C.c_int32, [C.c_uint64, C.c_uint64, C.c_int32, C.c_int32, C.c_int32])


def third_vertex(seed, s, n, a, b):
    """Third vertex of a 3-vertex move built on the pair (a, b), step s."""
    return int(_o.or_third_vertex(seed, s, n, a, b))


def p_rot3_fraction(p):
    """Probability of a 3-vertex move as the integer threshold on 32 random bits."""
    return min(1 << 32, max(0, int(round(p * 4294967296.0))))


def anneal_chain(o: "Oracle", pi, score, q, s0, steps, beta0, lam, p_rot3, state=None, pib=None, pool=None,
                 p_pool=0):
    """Replays steps s0..s0+steps-1 of one annealing chain on the CPU.
    state = [score, best, accepted, at_best]; returns (pi, pib, state) with pib
    the latest matching reaching the best score (pi itself when at_best).
    pool (k x 2 pairs) and p_pool (threshold on 32 random bits) give the
    proposal pool the chain draws from."""
    pi = _a32(pi).copy()
    pib = pi.copy() if pib is None else _a32(pib).copy()
    st = np.array(state if state is not None else [score, score, 0, 1], np.int64)
    pk = 0 if pool is None else int(np.asarray(pool).reshape(-1, 2).shape[0])
    pool = _a32(np.zeros((1, 2)) if pool is None else pool).reshape(-1, 2)
    _o.or_anneal_chain(o.h, _p32(pi), _p32(pib), q, s0, steps, beta0, lam, p_rot3, _p64(st), _p32(pool), pk,
                       int(p_pool))
    return pi, (pi.copy() if st[3] else pib), st


_sig(_o, "or_twohop_set", C.c_int32, [vp, C.c_int32, i32p])
_sig(_o, "or_twohop_pairs", None, [vp, i32p, C.c_int32, C.c_int32, C.c_uint64, i32p])


def twohop_set(o: "Oracle", v):
    out = np.empty(o.n + 1, np.int32)
    c = _o.or_twohop_set(o.h, v, _p32(out))
    return out[:c].copy()


def twohop_pairs(o: "Oracle", per_anchor, seed, anchors=None, n_anchors=None):
    if anchors is not None:
        a = _a32(anchors)
        na = a.shape[0]
        ap = _p32(a)
    else:
        na, ap = n_anchors, None
    pairs = np.empty((na * per_anchor, 2), np.int32)
    _o.or_twohop_pairs(o.h, ap, na, per_anchor, seed, _p32(pairs))
    return pairs


_sig(_o, "or_umajor_pairs", C.c_int, [C.c_uint64, C.c_int64, C.c_int32, i32p])


def umajor_pairs(seed, k, n):
    """The u-major uniform pair stream (k pairs, grouped by first vertex)."""
    uv = np.empty((k, 2), np.int32)
    _o.or_umajor_pairs(seed, k, n, _p32(uv))
    return uv
_sig(_o, "or_umajor_pairs_at", None, [C.c_uint64, C.c_int64, C.c_int32, i64p, C.c_int64, i32p])


_sig(_o, "or_amajor_triples", C.c_int, [C.c_uint64, C.c_int64, C.c_int32, i32p])
_sig(_o, "or_amajor_seed", C.c_uint64, [C.c_uint64])


def amajor_triples(seed, k, n):
    """The a-major uniform triple stream (k triples, grouped by first vertex)."""
    abc = np.empty((k, 3), np.int32)
    _o.or_amajor_triples(seed, k, n, _p32(abc))
    return abc


def amajor_seed(seed):
    """Seed of the a-major stream's first-vertex draws (a u-major stream seed)."""
    return int(_o.or_amajor_seed(seed))


def umajor_pairs_at(seed, k, n, idx):
    """Pairs at entry indices idx of the u-major stream (seed, k)."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    uv = np.empty((idx.shape[0], 2), np.int32)
    _o.or_umajor_pairs_at(seed, k, n, _p64(idx), idx.shape[0], _p32(uv))
    return uv
