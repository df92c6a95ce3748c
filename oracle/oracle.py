"""ctypes access to the oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a plain-C restatement of minipoa-rs's POA path
(src/align.rs, src/graph.rs, src/lib.rs). Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg import this module, and only as the checker.

`params` arguments are duck-typed: any object with the attributes of
src/lib.rs:47-71 `Params` (match_score, mismatch, gap_open, gap_ext, b, f,
adaptive, seeding, k, w, progressive, threads, win_band, win_max,
mm_filter_ratio, max_occ, tr_dist, result, verbose, alphabet, matrix,
is_rna) plus an optional `scalar_dp` flag.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from collections import namedtuple
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_poa.so"
# Mutation tests load a mutant build of the same source (ORACLE_MUTANT=k)
# through this variable; everything else uses the default build.
if os.environ.get("ORACLE_LIB"):
    LIB_PATH = Path(os.environ["ORACLE_LIB"])

AlignedBase = namedtuple("AlignedBase", "node base aligned")


class OracleError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [
        ("match_score", C.c_int32), ("mismatch", C.c_int32),
        ("gap_open", C.c_int32), ("gap_ext", C.c_int32),
        ("b", C.c_uint64), ("f", C.c_uint64),
        ("adaptive", C.c_int32), ("seeding", C.c_int32),
        ("k", C.c_uint64), ("w", C.c_uint64),
        ("progressive", C.c_int32),
        ("threads", C.c_uint64), ("win_band", C.c_uint64), ("win_max", C.c_uint64),
        ("mm_filter_ratio", C.c_double),
        ("max_occ", C.c_uint64), ("tr_dist", C.c_uint64),
        ("result", C.c_uint8), ("verbose", C.c_uint8),
        ("is_rna", C.c_int32), ("scalar_dp", C.c_int32),
        ("alphabet_len", C.c_uint32),
        ("alphabet", C.POINTER(C.c_uint8)),
        ("matrix", C.POINTER(C.c_int32)),
    ]


class _Base(C.Structure):
    _fields_ = [("node", C.c_int64), ("base", C.c_int64), ("aligned", C.c_int64)]


class _Anchor(C.Structure):
    _fields_ = [("target_end", C.c_uint64), ("query_end", C.c_uint64), ("span", C.c_uint64),
                ("adaptive_before", C.c_int64)]


# src/minimizer.rs:11-26
Minimizer = namedtuple("Minimizer", "value end span")
Anchor = namedtuple("Anchor", "target_end query_end span adaptive_before")


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER(_Params)
        G = C.c_void_p
        B = C.POINTER(_Base)
        sz = C.c_size_t
        L.o_last_error.restype = C.c_char_p
        L.og_new.argtypes = [P]; L.og_new.restype = G
        L.og_free.argtypes = [G]
        u64p_ = C.POINTER(C.c_uint64)
        L.og_from_arrays.argtypes = [P, sz] + [u64p_] * 9 + [C.POINTER(C.c_int64), sz, u64p_,
                                                             sz, u64p_, u64p_, u64p_]
        L.og_from_arrays.restype = G
        L.og_add_node.argtypes = [G, sz]; L.og_add_node.restype = sz
        L.og_add_edge.argtypes = [G, sz, sz]
        L.og_add_path.argtypes = [G, sz, B, sz]; L.og_add_path.restype = C.c_int
        L.og_toposort.argtypes = [G, C.c_int]; L.og_toposort.restype = C.c_int
        L.og_n_nodes.argtypes = [G]; L.og_n_nodes.restype = sz
        L.og_node_info.argtypes = [G, sz, C.POINTER(C.c_uint64)]
        L.og_node_aligned.argtypes = [G, sz, sz]; L.og_node_aligned.restype = C.c_int64
        L.og_node_edges.argtypes = [G, sz] + [C.POINTER(C.c_uint64)] * 4
        L.og_rank_len.argtypes = [G]; L.og_rank_len.restype = sz
        L.og_rank.argtypes = [G, C.POINTER(C.c_uint64)]
        L.og_n_paths.argtypes = [G]; L.og_n_paths.restype = sz
        L.og_consensus.argtypes = [G, P, C.POINTER(C.c_uint8), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64), sz]
        L.og_consensus.restype = sz
        L.og_msa.argtypes = [G, P, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), sz,
                             C.POINTER(C.c_void_p)]
        L.og_msa.restype = C.c_int
        L.og_gfa.argtypes = [G, P, C.POINTER(C.c_char_p), sz, C.POINTER(C.c_void_p)]
        L.og_gfa.restype = C.c_int
        L.o_free.argtypes = [C.c_void_p]
        L.o_encode.argtypes = [P, C.c_uint8]; L.o_encode.restype = sz
        L.o_validate.argtypes = [P]; L.o_validate.restype = C.c_int
        L.o_insertion_prefix.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                         C.c_int32, C.POINTER(C.c_int32)]
        L.o_i32_scores_safe.argtypes = [P, G, sz, sz, sz]; L.o_i32_scores_safe.restype = C.c_int
        for name in ("o_poa", "o_poa_inner", "o_poa_simd_i32", "o_poa_scalar_i64"):
            fn = getattr(L, name)
            fn.argtypes = [P, G, sz, sz, C.c_char_p, sz, C.c_int, sz, B, sz, C.POINTER(sz)]
            fn.restype = C.c_int
        L.o_align.argtypes = [P, G, C.c_char_p, sz, B, sz, C.POINTER(sz)]
        L.o_align.restype = C.c_int
        L.o_cells.restype = C.c_uint64
        L.o_last_score.restype = C.c_int64
        L.o_arena_capacity.restype = sz
        L.o_align_sequences.argtypes = [P, C.POINTER(C.c_char_p), C.POINTER(sz), sz]
        L.o_align_sequences.restype = G
        L.o_extend_graph.argtypes = [P, G, C.POINTER(C.c_char_p), C.POINTER(sz), sz, sz]
        L.o_extend_graph.restype = C.c_int
        L.o_batch_run.argtypes = [P, sz, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]
        L.o_batch_run.restype = C.c_uint64
        L.o_batch_run_gfa.argtypes = [P, sz, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.c_void_p, C.c_int, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]
        L.o_batch_run_gfa.restype = C.c_uint64
        u64p = C.POINTER(C.c_uint64)
        A = C.POINTER(_Anchor)
        L.o_sketch.argtypes = [C.c_char_p, sz, sz, sz, u64p, u64p, sz]
        L.o_sketch.restype = sz
        L.o_repeat_intervals.argtypes = [u64p, u64p, u64p, sz, sz, sz, u64p, u64p, sz]
        L.o_repeat_intervals.restype = sz
        L.o_collect_anchors.argtypes = [P, C.c_char_p, sz, C.c_char_p, sz, A, sz]
        L.o_collect_anchors.restype = sz
        L.o_chain.argtypes = [A, sz, sz, A, sz]
        L.o_chain.restype = sz
        L.o_partition_windows.argtypes = [A, sz, sz, sz, sz, sz, A, sz, C.POINTER(C.c_int)]
        L.o_partition_windows.restype = sz
        L.o_anchor_windows.argtypes = [P, C.c_char_p, sz, C.c_char_p, sz, A, sz,
                                       C.POINTER(C.c_int)]
        L.o_anchor_windows.restype = sz
        L.o_guide_order.argtypes = [P, C.POINTER(C.c_char_p), C.POINTER(sz), sz, C.POINTER(sz)]
        L.o_guide_order.restype = None
        _lib = L
    return _lib


def _err():
    return OracleError(lib().o_last_error().decode())


class _P:
    """Keeps the ctypes buffers of one params conversion alive."""

    def __init__(self, p):
        alphabet = bytes(p.alphabet)
        matrix = list(p.matrix)
        self.alpha = (C.c_uint8 * max(1, len(alphabet)))(*alphabet)
        self.mat = (C.c_int32 * max(1, len(matrix)))(*matrix)
        valid_dims = len(matrix) == len(alphabet) * len(alphabet)
        self.s = _Params(
            int(p.match_score), int(p.mismatch), int(p.gap_open), int(p.gap_ext),
            int(p.b), int(p.f), int(bool(p.adaptive)), int(bool(p.seeding)),
            int(p.k), int(p.w), int(bool(p.progressive)),
            int(p.threads), int(p.win_band), int(p.win_max), float(p.mm_filter_ratio),
            int(p.max_occ), int(p.tr_dist), int(p.result), int(p.verbose),
            int(bool(getattr(p, "is_rna", False))), int(bool(getattr(p, "scalar_dp", False))),
            len(alphabet) if valid_dims else 0,
            C.cast(self.alpha, C.POINTER(C.c_uint8)),
            C.cast(self.mat, C.POINTER(C.c_int32)) if valid_dims and matrix else None,
        )

    @property
    def ref(self):
        return C.byref(self.s)


def encode(params, base):
    return lib().o_encode(_P(params).ref, base if isinstance(base, int) else ord(base))


def validate(params):
    if lib().o_validate(_P(params).ref):
        raise _err()


def graph_from_arrays(params, nodes, aligned, order, paths):
    """A graph copied from plain data: nodes = list of dicts with rank,
    parent, base, incoming, in_weights, outgoing, out_weights; aligned[i] =
    list of alphabet_len entries (None or node id); order = rank order;
    paths = [(seq_id, [node ids])]."""
    import numpy as np

    def u64(a):
        return np.ascontiguousarray(np.array(a, dtype=np.uint64).reshape(-1))

    n = len(nodes)
    cols = {}
    for key in ("rank", "parent", "base"):
        cols[key] = u64([nd[key] for nd in nodes])
    in_off = u64([0] + list(np.cumsum([len(nd["incoming"]) for nd in nodes])))
    out_off = u64([0] + list(np.cumsum([len(nd["outgoing"]) for nd in nodes])))
    in_node = u64([x for nd in nodes for x in nd["incoming"]] or [0])
    in_w = u64([x for nd in nodes for x in nd["in_weights"]] or [0])
    out_node = u64([x for nd in nodes for x in nd["outgoing"]] or [0])
    out_w = u64([x for nd in nodes for x in nd["out_weights"]] or [0])
    al = np.ascontiguousarray(np.array([[-1 if x is None else x for x in row] for row in aligned],
                                       dtype=np.int64).reshape(-1))
    order_a = u64(order or [0])
    pseq = u64([s for s, _ in paths] or [0])
    poff = u64([0] + list(np.cumsum([len(ids) for _, ids in paths])))
    pids = u64([x for _, ids in paths for x in ids] or [0])
    p64 = C.POINTER(C.c_uint64)
    ptr = lambda a: a.ctypes.data_as(p64)
    h = lib().og_from_arrays(
        _P(params).ref, n, ptr(cols["rank"]), ptr(cols["parent"]), ptr(cols["base"]),
        ptr(in_off), ptr(in_node), ptr(in_w), ptr(out_off), ptr(out_node), ptr(out_w),
        al.ctypes.data_as(C.POINTER(C.c_int64)), len(order), ptr(order_a), len(paths), ptr(pseq),
        ptr(poff), ptr(pids))
    return Graph(_handle=h)


class Graph:
    """src/graph.rs:20-331 through the oracle."""

    def __init__(self, params=None, _handle=None):
        self._h = _handle if _handle is not None else lib().og_new(_P(params).ref)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().og_free(self._h)
            self._h = None

    def add_node(self, base):
        return lib().og_add_node(self._h, base)

    def add_edge(self, a, b):
        lib().og_add_edge(self._h, a, b)

    def add_path(self, seq_id, path):
        arr = (_Base * max(1, len(path)))()
        for k, it in enumerate(path):
            arr[k] = _Base(-1 if it[0] is None else it[0], it[1], -1 if it[2] is None else it[2])
        if lib().og_add_path(self._h, seq_id, arr, len(path)):
            raise _err()

    def toposort(self, grouped=False):
        if lib().og_toposort(self._h, int(grouped)):
            raise _err()

    @property
    def n_nodes(self):
        return lib().og_n_nodes(self._h)

    def node(self, i):
        out = (C.c_uint64 * 6)()
        lib().og_node_info(self._h, i, out)
        rank, parent, base, coverage, nin, nout = list(out)
        a = (C.c_uint64 * max(1, nin))(); aw = (C.c_uint64 * max(1, nin))()
        b = (C.c_uint64 * max(1, nout))(); bw = (C.c_uint64 * max(1, nout))()
        lib().og_node_edges(self._h, i, a, aw, b, bw)
        return dict(id=i, rank=rank, parent=parent, base=base, coverage=coverage,
                    incoming=list(a)[:nin], in_weights=list(aw)[:nin],
                    outgoing=list(b)[:nout], out_weights=list(bw)[:nout])

    def aligned(self, i, base):
        v = lib().og_node_aligned(self._h, i, base)
        return None if v < 0 else v

    @property
    def rank(self):
        n = lib().og_rank_len(self._h)
        out = (C.c_uint64 * max(1, n))()
        lib().og_rank(self._h, out)
        return list(out)[:n]

    def consensus(self, params):
        n = self.n_nodes
        seq = (C.c_uint8 * n)(); ids = (C.c_uint64 * n)(); cov = (C.c_uint64 * n)()
        k = lib().og_consensus(self._h, _P(params).ref, seq, ids, cov, n)
        return bytes(seq)[:k], list(ids)[:k], list(cov)[:k]

    def msa(self, params, seqs):
        names = (C.c_char_p * max(1, len(seqs)))(*[s.name.encode() for s in seqs])
        comments = (C.c_char_p * max(1, len(seqs)))(*[s.comment.encode() for s in seqs])
        out = C.c_void_p()
        if lib().og_msa(self._h, _P(params).ref, names, comments, len(seqs), C.byref(out)):
            raise _err()
        text = C.string_at(out.value).decode()
        lib().o_free(out)
        return text

    def gfa(self, params, seqs):
        names = (C.c_char_p * max(1, len(seqs)))(*[s.name.encode() for s in seqs])
        out = C.c_void_p()
        if lib().og_gfa(self._h, _P(params).ref, names, len(seqs), C.byref(out)):
            raise _err()
        text = C.string_at(out.value).decode()
        lib().o_free(out)
        return text


def _path_call(fn, params, graph, *args, qlen_hint=0):
    cap = qlen_hint + 16
    while True:
        buf = (_Base * cap)()
        n = C.c_size_t(0)
        if fn(*args, buf, cap, C.byref(n)):
            raise _err()
        if n.value <= cap:
            return [AlignedBase(None if b.node < 0 else b.node, b.base,
                                None if b.aligned < 0 else b.aligned) for b in buf[:n.value]]
        cap = n.value + 16


def align(params, graph, query):
    q = bytes(query)
    P = _P(params)
    return _path_call(lib().o_align, params, graph, P.ref, graph._h, q, len(q), qlen_hint=len(q))


def _poa(name, params, graph, beg, end, query, adaptive, band):
    q = bytes(query)
    P = _P(params)
    return _path_call(getattr(lib(), name), params, graph, P.ref, graph._h, beg, end, q, len(q),
                      int(bool(adaptive)), band, qlen_hint=len(q))


def poa(params, graph, beg, end, query, adaptive, band):
    return _poa("o_poa", params, graph, beg, end, query, adaptive, band)


def poa_inner(params, graph, beg, end, query, adaptive, band):
    return _poa("o_poa_inner", params, graph, beg, end, query, adaptive, band)


def poa_simd_i32(params, graph, beg, end, query, adaptive, band):
    return _poa("o_poa_simd_i32", params, graph, beg, end, query, adaptive, band)


def poa_scalar_i64(params, graph, beg, end, query, adaptive, band):
    return _poa("o_poa_scalar_i64", params, graph, beg, end, query, adaptive, band)


def result_or_error(fn, *args):
    """Ok(path) / Err(message) as a comparable tuple, like Rust's Result."""
    try:
        return ("ok", fn(*args))
    except OracleError as e:
        return ("err", str(e))


def i32_scores_safe(params, graph, beg, end, qlen):
    return bool(lib().o_i32_scores_safe(_P(params).ref, graph._h, beg, end, qlen))


def insertion_prefix(md, carry, open_, ext):
    a = (C.c_int32 * 8)(*md)
    out = (C.c_int32 * 8)()
    lib().o_insertion_prefix(a, carry, open_, ext, out)
    return list(out)


def cells():
    return lib().o_cells()


def last_score():
    """M at (n-1, m-1) of the calling thread's last DP fill."""
    return lib().o_last_score()


def reset_cells():
    lib().o_reset_cells()


def arena_capacity():
    return lib().o_arena_capacity()


def arena_clear():
    lib().o_arena_clear()


def align_sequences(params, seqs):
    raw = [bytes(s.seq) for s in seqs]
    arr = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_size_t * max(1, len(raw)))(*[len(r) for r in raw])
    h = lib().o_align_sequences(_P(params).ref, arr, lens, len(raw))
    if not h:
        raise _err()
    return Graph(_handle=h)


def extend_graph(params, graph, seqs, existing):
    raw = [bytes(s.seq) for s in seqs]
    arr = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_size_t * max(1, len(raw)))(*[len(r) for r in raw])
    if lib().o_extend_graph(_P(params).ref, graph._h, arr, lens, len(raw), existing):
        raise _err()


def batch_run(params, job_off, read_off, bases, nthreads, gfa=False):
    """Job-parallel progressive POA + consensus/MSA hash per job (numpy
    arrays). With gfa: also the FNV-1a 64 of each job's GFA text, returned
    as a third value."""
    import numpy as np

    job_off = np.ascontiguousarray(job_off, dtype=np.uint64)
    read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    njobs = len(job_off) - 1
    out = np.zeros(njobs, dtype=np.uint64)
    outg = np.zeros(njobs, dtype=np.uint64)
    u64p = C.POINTER(C.c_uint64)
    cells_ = lib().o_batch_run_gfa(
        _P(params).ref, njobs, job_off.ctypes.data_as(u64p), read_off.ctypes.data_as(u64p),
        bases.ctypes.data_as(C.c_void_p), int(nthreads), out.ctypes.data_as(u64p),
        outg.ctypes.data_as(u64p) if gfa else None)
    return (int(cells_), out, outg) if gfa else (int(cells_), out)


# ---- minimizer seeding (src/minimizer.rs:32-488) ----

def _grow(call, cap=64):
    """Calls call(cap) -> (count, result) until the buffer held everything."""
    while True:
        n, res = call(cap)
        if n <= cap:
            return res(n)
        cap = n + 16


def sketch(seq, k, w):
    """src/minimizer.rs:32-76 (positions parity unpinned)."""
    s = bytes(seq)

    def call(cap):
        v = (C.c_uint64 * cap)()
        e = (C.c_uint64 * cap)()
        n = lib().o_sketch(s, len(s), k, w, v, e, cap)
        return n, lambda n: [Minimizer(v[i], e[i], k) for i in range(n)]
    return _grow(call, len(s) + 1)


def repeat_intervals(mm, length, max_dist):
    """src/minimizer.rs:78-110"""
    n = len(mm)
    v = (C.c_uint64 * max(1, n))(*[m.value for m in mm])
    e = (C.c_uint64 * max(1, n))(*[m.end for m in mm])
    sp = (C.c_uint64 * max(1, n))(*[m.span for m in mm])

    def call(cap):
        a = (C.c_uint64 * cap)()
        b = (C.c_uint64 * cap)()
        r = lib().o_repeat_intervals(v, e, sp, n, length, max_dist, a, b, cap)
        return r, lambda r: [(a[i], b[i]) for i in range(r)]
    return _grow(call, n + 1)


def _anchors_in(anchors):
    arr = (_Anchor * max(1, len(anchors)))()
    for i, a in enumerate(anchors):
        arr[i] = _Anchor(a.target_end, a.query_end, a.span, int(bool(a.adaptive_before)))
    return arr


def _anchors_out(arr, n):
    return [Anchor(arr[i].target_end, arr[i].query_end, arr[i].span, bool(arr[i].adaptive_before))
            for i in range(n)]


def collect_anchors(params, query, target):
    """src/minimizer.rs:128-176"""
    P = _P(params)
    q, t = bytes(query), bytes(target)

    def call(cap):
        out = (_Anchor * cap)()
        n = lib().o_collect_anchors(P.ref, q, len(q), t, len(t), out, cap)
        return n, lambda n: _anchors_out(out, n)
    return _grow(call)


def chain(anchors, k):
    """src/minimizer.rs:199-377"""
    arr = _anchors_in(anchors)

    def call(cap):
        out = (_Anchor * cap)()
        n = lib().o_chain(arr, len(anchors), k, out, cap)
        return n, lambda n: _anchors_out(out, n)
    return _grow(call, len(anchors) + 1)


def partition_windows(chain_, tlen, qlen, max_drift, max_win):
    """src/minimizer.rs:379-467 -> (boundaries, tail_adaptive)"""
    arr = _anchors_in(chain_)
    tail = C.c_int(0)

    def call(cap):
        out = (_Anchor * cap)()
        n = lib().o_partition_windows(arr, len(chain_), tlen, qlen, max_drift, max_win, out, cap,
                                      C.byref(tail))
        return n, lambda n: _anchors_out(out, n)
    res = _grow(call, len(chain_) + 1)
    return res, bool(tail.value)


def anchor_windows(params, query, consensus):
    """src/minimizer.rs:475-488 -> (boundaries, tail_adaptive)"""
    P = _P(params)
    q, t = bytes(query), bytes(consensus)
    tail = C.c_int(0)

    def call(cap):
        out = (_Anchor * cap)()
        n = lib().o_anchor_windows(P.ref, q, len(q), t, len(t), out, cap, C.byref(tail))
        return n, lambda n: _anchors_out(out, n)
    res = _grow(call)
    return res, bool(tail.value)


def guide_order(params, seqs):
    """src/minimizer.rs:489-578 (seqs: Sequence records or byte strings)."""
    raw = [bytes(getattr(s, "seq", s)) for s in seqs]
    P = _P(params)
    arr = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_size_t * max(1, len(raw)))(*[len(r) for r in raw])
    out = (C.c_size_t * max(1, len(raw)))()
    lib().o_guide_order(P.ref, arr, lens, len(raw), out)
    return list(out)[:len(raw)]
