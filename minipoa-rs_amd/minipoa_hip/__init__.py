"""ctypes binding of libminipoa_hip.so (include/minipoa_hip.h).

Mirrors minipoa-rs's library surface for the DP path: `poa`/`poa_inner`
(src/align.rs:192-211, 531-545), `align` (81-98), `align_sequences`
(src/lib.rs:177-204), the graph (src/graph.rs) and the batched engine.
There is no CPU fallback: every DP call runs the gfx950 kernels, and a missing
library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from collections import namedtuple
from pathlib import Path

from .params import AlignedBase, Params, Sequence, default_matrix, records

PKG = Path(__file__).resolve().parent.parent
LIB_PATH = PKG / "libminipoa_hip.so"
# MPOA_LIB names another build of the library (for side-by-side comparisons
# of build options); the default is the in-tree build.
if os.environ.get("MPOA_LIB"):
    LIB_PATH = Path(os.environ["MPOA_LIB"])

STATUS = {
    0: "MPOA_OK", 1: "MPOA_ERR_ANCHOR_NOT_IN_GRAPH", 2: "MPOA_ERR_ANCHORS_NOT_ORDERED",
    3: "MPOA_ERR_COORDINATE_LIMITS", 4: "MPOA_ERR_NO_PATH", 5: "MPOA_ERR_BAND_NO_INTERSECT",
    6: "MPOA_ERR_BAND_EXCLUDES_SOURCE", 7: "MPOA_ERR_TRACEBACK_LEFT_BAND",
    8: "MPOA_ERR_TRACEBACK_NO_ADVANCE", 9: "MPOA_ERR_TRACEBACK_NO_PREDECESSOR",
    10: "MPOA_ERR_ALLOC", 11: "MPOA_ERR_INVALID_PARAMS", 12: "MPOA_ERR_GRAPH",
    13: "MPOA_ERR_DEVICE", 14: "MPOA_ERR_NO_DEVICE", 15: "MPOA_ERR_UNSUPPORTED",
    16: "MPOA_ERR_BUFFER_TOO_SMALL", 17: "MPOA_ERR_INVALID_ARGUMENT", 18: "MPOA_ERR_PENDING",
    19: "MPOA_ERR_SEED_SPAN", 20: "MPOA_ERR_SCRATCH_LIMIT",
}
BUFFER_TOO_SMALL = 16
INVALID_ARGUMENT = 17
SCRATCH_LIMIT = 20
LAYOUT_AUTO, LAYOUT_VALUES, LAYOUT_RECORDS = 0, 1, 2  # mpoa_ctx_set_layout


class HipPoaError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


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


U64P = C.POINTER(C.c_uint64)


class _View(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint64), ("alphabet_len", C.c_uint64),
        ("rank", U64P), ("parent", U64P), ("base", U64P),
        ("in_off", U64P), ("in_node", U64P), ("in_weight", U64P),
        ("out_off", U64P), ("out_node", U64P), ("out_weight", U64P),
        ("aligned", C.POINTER(C.c_int64)),
        ("n_order", C.c_uint64), ("order", U64P),
    ]


class _Stats(C.Structure):
    _fields_ = [
        ("cells", C.c_uint64), ("dp_calls", C.c_uint64), ("retries", C.c_uint64),
        ("i64_calls", C.c_uint64), ("pred_cells", C.c_uint64), ("fill_launches", C.c_uint64),
        ("device_ms", C.c_double), ("fill_ms", C.c_double),
        ("host_ms", C.c_double), ("wall_ms", C.c_double),
        ("plan_ms", C.c_double), ("consume_ms", C.c_double), ("fuse_ms", C.c_double),
        ("render_ms", C.c_double), ("pack_ms", C.c_double), ("arena_retries", C.c_uint64),
        ("arena_bytes", C.c_uint64), ("record_launches", C.c_uint64),
        ("resident_jobs", C.c_uint64), ("evicted_jobs", C.c_uint64), ("step_replays", C.c_uint64),
        ("buckets", C.c_uint64), ("deferred", C.c_uint64), ("resident_ms", C.c_double),
        ("resident_device_ms", C.c_double),
    ]


class _ServerOpts(C.Structure):
    _fields_ = [("admit_bytes", C.c_uint64), ("slots", C.c_uint32), ("render", C.c_int32),
                ("resident", C.c_int32)]


class _JobTimes(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double),
                ("cells", C.c_uint64)]


class _Anchor(C.Structure):
    _fields_ = [("target_end", C.c_uint64), ("query_end", C.c_uint64), ("span", C.c_uint64),
                ("adaptive_before", C.c_int64)]


# src/minimizer.rs:11-26
Minimizer = namedtuple("Minimizer", "value end span")
Anchor = namedtuple("Anchor", "target_end query_end span adaptive_before")


# Exported symbols of include/minipoa_hip.h with their ctypes signatures.
_P = C.POINTER(_Params)
_V = C.POINTER(_View)
_B = C.POINTER(_Base)
_u64 = C.c_uint64
_vp = C.c_void_p
SIGNATURES = {
    "mpoa_backend_name": (C.c_char_p, []),
    "mpoa_last_error": (C.c_char_p, []),
    "mpoa_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "mpoa_ctx_destroy": (None, [_vp]),
    "mpoa_ctx_set_layout": (C.c_int, [_vp, C.c_int]),
    "mpoa_ctx_device_bytes": (_u64, [_vp]),
    "mpoa_poa": (C.c_int, [_vp, _P, _V, _u64, _u64, C.c_char_p, _u64, C.c_int, _u64, _B, _u64,
                           U64P, C.POINTER(C.c_int64)]),
    "mpoa_poa_inner": (C.c_int, [_vp, _P, _V, _u64, _u64, C.c_char_p, _u64, C.c_int, _u64, _B,
                                 _u64, U64P, C.POINTER(C.c_int64)]),
    "mpoa_poa_noalloc": (C.c_int, [_vp, _P, _V, _u64, _u64, C.c_char_p, _u64, C.c_int, _u64, _B,
                                   _u64, U64P, C.POINTER(C.c_int64)]),
    "mpoa_poa_kernel": (C.c_int, [_vp, _P, _V, _u64, _u64, C.c_char_p, _u64, C.c_int, _u64,
                                  C.c_int, _B, _u64, U64P, C.POINTER(C.c_int64)]),
    "mpoa_i32_scores_safe": (C.c_int, [_P, _V, _u64, _u64, _u64]),
    "mpoa_align": (C.c_int, [_vp, _P, _V, C.c_char_p, _u64, _B, _u64, U64P]),
    "mpoa_graph_new": (_vp, [_P]),
    "mpoa_graph_free": (None, [_vp]),
    "mpoa_graph_add_node": (_u64, [_vp, _u64]),
    "mpoa_graph_add_edge": (C.c_int, [_vp, _u64, _u64]),
    "mpoa_graph_add_path": (C.c_int, [_vp, _u64, _B, _u64]),
    "mpoa_graph_toposort": (C.c_int, [_vp, C.c_int]),
    "mpoa_graph_view_of": (C.c_int, [_vp, _V]),
    "mpoa_graph_n_paths": (_u64, [_vp]),
    "mpoa_graph_path": (C.c_int, [_vp, _u64, U64P, U64P, _u64, U64P]),
    "mpoa_graph_consensus": (_u64, [_vp, _P, C.POINTER(C.c_uint8), U64P, U64P, _u64]),
    "mpoa_graph_msa": (C.c_int, [_vp, _P, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), _u64,
                                 C.c_char_p, _u64, U64P]),
    "mpoa_graph_gfa": (C.c_int, [_vp, _P, C.POINTER(C.c_char_p), _u64, C.c_char_p, _u64, U64P]),
    "mpoa_params_validate": (C.c_int, [_P]),
    "mpoa_read_sequences": (C.c_int, [C.c_char_p, _u64, C.POINTER(_vp)]),
    "mpoa_records_len": (_u64, [_vp]),
    "mpoa_record_get": (C.c_int, [_vp, _u64, C.POINTER(C.c_void_p), U64P, C.POINTER(C.c_void_p),
                                  U64P, C.POINTER(C.c_void_p), U64P, C.POINTER(C.c_void_p), U64P]),
    "mpoa_records_free": (None, [_vp]),
    "mpoa_read_gfa": (C.c_int, [_P, C.c_char_p, _u64, C.POINTER(_vp), C.POINTER(_vp)]),
    "mpoa_sketch": (C.c_int, [C.c_char_p, _u64, _u64, _u64, U64P, U64P, _u64, U64P]),
    "mpoa_collect_anchors": (C.c_int, [_P, C.c_char_p, _u64, C.c_char_p, _u64, C.POINTER(_Anchor),
                                       _u64, U64P]),
    "mpoa_chain": (C.c_int, [C.POINTER(_Anchor), _u64, _u64, C.POINTER(_Anchor), _u64, U64P]),
    "mpoa_partition_windows": (C.c_int, [C.POINTER(_Anchor), _u64, _u64, _u64, _u64, _u64,
                                         C.POINTER(_Anchor), _u64, U64P, C.POINTER(C.c_int)]),
    "mpoa_guide_order": (C.c_int, [_P, C.POINTER(C.c_char_p), U64P, _u64, U64P]),
    "mpoa_anchor_windows": (C.c_int, [_P, C.c_char_p, _u64, C.c_char_p, _u64, C.POINTER(_Anchor),
                                      _u64, U64P, C.POINTER(C.c_int)]),
    "mpoa_load_matrix": (C.c_int, [C.c_char_p, _u64, C.POINTER(C.c_uint8), C.POINTER(C.c_int32),
                                   C.c_uint32, C.POINTER(C.c_uint32)]),
    "mpoa_align_sequences": (C.c_int, [_vp, _P, C.POINTER(C.c_char_p), U64P, _u64,
                                       C.POINTER(_vp)]),
    "mpoa_extend_graph": (C.c_int, [_vp, _P, _vp, C.POINTER(C.c_char_p), U64P, _u64, _u64]),
    "mpoa_batch_run": (C.c_int, [_vp, _P, _u64, U64P, U64P, _vp, C.POINTER(_vp)]),
    "mpoa_batch_run_resident": (C.c_int, [_vp, _P, _u64, U64P, U64P, _vp, C.POINTER(_vp)]),
    "mpoa_resident_check": (C.c_int, [_P, _vp, U64P, _u64, _B, U64P, C.c_uint32, C.c_char_p, _u64]),
    "mpoa_batch_free": (None, [_vp]),
    "mpoa_batch_stats_get": (C.c_int, [_vp, C.POINTER(_Stats)]),
    "mpoa_batch_job_status": (C.c_int, [_vp, _u64]),
    "mpoa_batch_job_output": (C.c_int, [_vp, _u64, C.c_int, C.c_char_p, _u64, U64P]),
    "mpoa_server_create": (C.c_int, [_vp, _P, C.POINTER(_ServerOpts), C.POINTER(_vp)]),
    "mpoa_server_submit": (C.c_int, [_vp, _vp, U64P, _u64, _vp, _u64, U64P]),
    "mpoa_server_done": (C.c_int, [_vp, _u64]),
    "mpoa_server_wait": (C.c_int, [_vp, _u64]),
    "mpoa_server_job_status": (C.c_int, [_vp, _u64]),
    "mpoa_server_job_output": (C.c_int, [_vp, _u64, C.c_int, C.c_char_p, _u64, U64P]),
    "mpoa_server_job_times": (C.c_int, [_vp, _u64, C.POINTER(_JobTimes)]),
    "mpoa_server_job_graph": (C.c_int, [_vp, _u64, C.POINTER(_vp)]),
    "mpoa_server_stats_get": (C.c_int, [_vp, C.POINTER(_Stats)]),
    "mpoa_server_now_ms": (C.c_double, [_vp]),
    "mpoa_server_free": (None, [_vp]),
}

_lib = None


def lib():
    """Load the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HipPoaError(14, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().mpoa_last_error().decode(errors="replace")


def _check(rc):
    if rc != 0:
        raise HipPoaError(rc, last_error())


def backend_name():
    return lib().mpoa_backend_name().decode()


class CParams:
    """Keeps the ctypes buffers of one Params conversion alive."""

    def __init__(self, p: Params):
        alphabet = bytes(p.alphabet)
        matrix = list(p.matrix)
        self.alpha = (C.c_uint8 * max(1, len(alphabet)))(*alphabet)
        self.mat = (C.c_int32 * max(1, len(matrix)))(*matrix)
        ok = len(matrix) == len(alphabet) ** 2 and len(matrix) > 0
        self.s = _Params(
            int(p.match_score), int(p.mismatch), int(p.gap_open), int(p.gap_ext),
            int(p.b), int(p.f), int(bool(p.adaptive)), int(bool(p.seeding)),
            int(p.k), int(p.w), int(bool(p.progressive)),
            int(p.threads), int(p.win_band), int(p.win_max), float(p.mm_filter_ratio),
            int(p.max_occ), int(p.tr_dist), int(p.result), int(p.verbose),
            int(bool(p.is_rna)), int(bool(getattr(p, "scalar_dp", False))),
            len(alphabet) if ok else 0,
            C.cast(self.alpha, C.POINTER(C.c_uint8)),
            C.cast(self.mat, C.POINTER(C.c_int32)) if ok else None,
        )

    @property
    def ref(self):
        return C.byref(self.s)


class Context:
    """One device context (stream, device arenas, staging, host workers)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().mpoa_ctx_create(device, C.byref(h)))
        self._h = h

    def set_layout(self, layout):
        """DP layout in device memory: LAYOUT_AUTO, LAYOUT_VALUES or
        LAYOUT_RECORDS (mpoa_ctx_set_layout); results are the same."""
        _check(lib().mpoa_ctx_set_layout(self._h, int(layout)))

    def close(self):
        if getattr(self, "_h", None):
            lib().mpoa_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def device_bytes(self):
        return lib().mpoa_ctx_device_bytes(self._h)


class GraphView:
    """A flat mpoa_graph_view built from node records (keeps arrays alive).

    nodes: list of dicts with rank, parent, base, incoming, in_weights,
    outgoing, out_weights; aligned: n x alphabet list (None entries allowed);
    order: graph.rank (rank -> id)."""

    def __init__(self, nodes, aligned, order, alphabet_len):
        n = len(nodes)
        u64 = C.c_uint64

        def arr(vals, t=u64):
            return (t * max(1, len(vals)))(*vals)

        in_off, out_off, ins, inw, outs, outw = [0], [0], [], [], [], []
        for nd in nodes:
            ins += nd["incoming"]
            inw += nd["in_weights"]
            outs += nd["outgoing"]
            outw += nd["out_weights"]
            in_off.append(len(ins))
            out_off.append(len(outs))
        flat_al = []
        for row in aligned:
            flat_al += [-1 if x is None else x for x in row]
        self._keep = [
            arr([nd["rank"] for nd in nodes]), arr([nd["parent"] for nd in nodes]),
            arr([nd["base"] for nd in nodes]), arr(in_off), arr(ins), arr(inw), arr(out_off),
            arr(outs), arr(outw), arr(flat_al, C.c_int64), arr(order),
        ]
        k = self._keep
        self.s = _View(n, alphabet_len, k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8],
                       k[9], len(order), k[10])

    @property
    def ref(self):
        return C.byref(self.s)


def view_from_graph(og, alphabet_len):
    """Flatten any graph object exposing node(i), aligned(i, b), rank and
    n_nodes (as graph::Graph does) into a GraphView."""
    nodes = [og.node(i) for i in range(og.n_nodes)]
    aligned = [[og.aligned(i, b) for b in range(alphabet_len)] for i in range(og.n_nodes)]
    return GraphView(nodes, aligned, og.rank, alphabet_len)


def _to_bases(buf, n):
    return [AlignedBase(None if b.node < 0 else b.node, b.base,
                        None if b.aligned < 0 else b.aligned) for b in buf[:n]]


def _path_call(fn, args, cap):
    while True:
        buf = (_Base * cap)()
        n = C.c_uint64(0)
        rc = fn(*args[0], buf, cap, C.byref(n), *args[1])
        if rc == BUFFER_TOO_SMALL:
            cap = n.value + 16
            continue
        _check(rc)
        return _to_bases(buf, n.value)


class Graph:
    """Library-owned graph (src/graph.rs semantics)."""

    def __init__(self, params: Params, _handle=None):
        self._cp = CParams(params)
        self._h = _handle if _handle is not None else lib().mpoa_graph_new(self._cp.ref)
        if not self._h:
            raise HipPoaError(11, last_error())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mpoa_graph_free(self._h)
            self._h = None

    def add_node(self, base):
        nid = lib().mpoa_graph_add_node(self._h, base)
        if nid == (1 << 64) - 1:
            raise HipPoaError(INVALID_ARGUMENT, last_error())
        return nid

    def add_edge(self, a, b):
        _check(lib().mpoa_graph_add_edge(self._h, a, b))

    def add_path(self, seq_id, path):
        arr = (_Base * max(1, len(path)))()
        for k, it in enumerate(path):
            arr[k] = _Base(-1 if it[0] is None else it[0], it[1], -1 if it[2] is None else it[2])
        _check(lib().mpoa_graph_add_path(self._h, seq_id, arr, len(path)))

    def toposort(self, grouped=False):
        _check(lib().mpoa_graph_toposort(self._h, int(grouped)))

    def view(self):
        v = _View()
        _check(lib().mpoa_graph_view_of(self._h, C.byref(v)))
        return v

    @property
    def n_nodes(self):
        return self.view().n_nodes

    def node(self, i):
        v = self.view()
        a, b = v.in_off[i], v.in_off[i + 1]
        c, d = v.out_off[i], v.out_off[i + 1]
        return dict(id=i, rank=v.rank[i], parent=v.parent[i], base=v.base[i],
                    incoming=list(v.in_node[a:b]), in_weights=list(v.in_weight[a:b]),
                    outgoing=list(v.out_node[c:d]), out_weights=list(v.out_weight[c:d]))

    def aligned(self, i, base):
        v = self.view()
        x = v.aligned[i * v.alphabet_len + base]
        return None if x < 0 else x

    def snapshot(self):
        """Every node (as node(i) gives it), the aligned table and the rank
        order from one view: (nodes, aligned, order)."""
        v = self.view()
        n, a = v.n_nodes, v.alphabet_len
        io, oo = v.in_off[:n + 1], v.out_off[:n + 1]
        ein, ew = v.in_node[:io[n]], v.in_weight[:io[n]]
        eout, eow = v.out_node[:oo[n]], v.out_weight[:oo[n]]
        rank, parent, base = v.rank[:n], v.parent[:n], v.base[:n]
        al = v.aligned[:n * a]
        nodes = [dict(id=i, rank=rank[i], parent=parent[i], base=base[i],
                      incoming=ein[io[i]:io[i + 1]], in_weights=ew[io[i]:io[i + 1]],
                      outgoing=eout[oo[i]:oo[i + 1]], out_weights=eow[oo[i]:oo[i + 1]])
                 for i in range(n)]
        aligned = [[None if x < 0 else x for x in al[i * a:(i + 1) * a]] for i in range(n)]
        return nodes, aligned, list(v.order[:v.n_order])

    @property
    def rank(self):
        v = self.view()
        return list(v.order[:v.n_order])

    @property
    def paths(self):
        """graph.paths: [(seq_id, [node ids])] in alignment order."""
        out = []
        L = lib()
        for k in range(L.mpoa_graph_n_paths(self._h)):
            sid, n = C.c_uint64(), C.c_uint64()
            rc = L.mpoa_graph_path(self._h, k, C.byref(sid), None, 0, C.byref(n))
            if rc not in (0, BUFFER_TOO_SMALL):
                _check(rc)
            ids = (C.c_uint64 * max(1, n.value))()
            _check(L.mpoa_graph_path(self._h, k, C.byref(sid), ids, n.value, C.byref(n)))
            out.append((sid.value, list(ids)[:n.value]))
        return out

    def consensus(self, params):
        n = self.n_nodes
        seq = (C.c_uint8 * max(1, n))()
        ids = (C.c_uint64 * max(1, n))()
        cov = (C.c_uint64 * max(1, n))()
        cp = CParams(params)
        k = lib().mpoa_graph_consensus(self._h, cp.ref, seq, ids, cov, n)
        return bytes(seq)[:k], list(ids)[:k], list(cov)[:k]

    def _text(self, fn, *args):
        cap = 1 << 16
        while True:
            buf = C.create_string_buffer(cap)
            n = C.c_uint64(0)
            rc = fn(*args, buf, cap, C.byref(n))
            if rc == BUFFER_TOO_SMALL:
                cap = n.value + 1
                continue
            _check(rc)
            return buf.raw[:n.value].decode()

    def msa(self, params, seqs):
        names = (C.c_char_p * max(1, len(seqs)))(*[s.name.encode() for s in seqs])
        comments = (C.c_char_p * max(1, len(seqs)))(*[s.comment.encode() for s in seqs])
        cp = CParams(params)
        return self._text(lib().mpoa_graph_msa, self._h, cp.ref, names, comments, len(seqs))

    def gfa(self, params, seqs):
        names = (C.c_char_p * max(1, len(seqs)))(*[s.name.encode() for s in seqs])
        cp = CParams(params)
        return self._text(lib().mpoa_graph_gfa, self._h, cp.ref, names, len(seqs))


def _as_view(graph):
    if isinstance(graph, GraphView):
        return graph.ref, graph
    v = graph.view()
    return C.byref(v), v


def poa(ctx, params, graph, beg, end, query, adaptive, band, score=False):
    """src/align.rs:192-211 on the device (banded, then full-DP retry)."""
    cp = CParams(params)
    vref, keep = _as_view(graph)
    q = bytes(query)
    sc = C.c_int64(0)
    out = _path_call(lib().mpoa_poa, ((ctx._h, cp.ref, vref, beg, end, q, len(q),
                                       int(bool(adaptive)), band), (C.byref(sc),)), len(q) + 16)
    return (out, sc.value) if score else out


def poa_inner(ctx, params, graph, beg, end, query, adaptive, band, score=False):
    """src/align.rs:531-545 on the device (guard picks i32 or i64)."""
    cp = CParams(params)
    vref, keep = _as_view(graph)
    q = bytes(query)
    sc = C.c_int64(0)
    out = _path_call(lib().mpoa_poa_inner, ((ctx._h, cp.ref, vref, beg, end, q, len(q),
                                             int(bool(adaptive)), band), (C.byref(sc),)),
                     len(q) + 16)
    return (out, sc.value) if score else out


def poa_noalloc(ctx, params, graph, beg, end, query, adaptive, band, score=False):
    """`poa` as mpoa_poa_noalloc runs it: the same DP with every device
    buffer taken from the library's fixed per-GPU scratch (no allocation)."""
    cp = CParams(params)
    vref, keep = _as_view(graph)
    q = bytes(query)
    sc = C.c_int64(0)
    out = _path_call(lib().mpoa_poa_noalloc, ((ctx._h, cp.ref, vref, beg, end, q, len(q),
                                               int(bool(adaptive)), band), (C.byref(sc),)),
                     len(q) + 16)
    return (out, sc.value) if score else out


def poa_kernel(ctx, params, graph, beg, end, query, adaptive, band, bits, score=False):
    """One device kernel forced: bits 32 (i32 fast path) or 64 (i64)."""
    cp = CParams(params)
    vref, keep = _as_view(graph)
    q = bytes(query)
    sc = C.c_int64(0)
    out = _path_call(lib().mpoa_poa_kernel, ((ctx._h, cp.ref, vref, beg, end, q, len(q),
                                              int(bool(adaptive)), band, bits), (C.byref(sc),)),
                     len(q) + 16)
    return (out, sc.value) if score else out


def i32_scores_safe(params, graph, beg, end, qlen):
    cp = CParams(params)
    vref, keep = _as_view(graph)
    return bool(lib().mpoa_i32_scores_safe(cp.ref, vref, beg, end, qlen))


def align(ctx, params, graph, query):
    """src/align.rs:81-98 on the device."""
    cp = CParams(params)
    vref, keep = _as_view(graph)
    q = bytes(query)
    return _path_call(lib().mpoa_align, ((ctx._h, cp.ref, vref, q, len(q)), ()), len(q) + 16)


def result_or_error(fn, *args, **kw):
    try:
        return ("ok", fn(*args, **kw))
    except HipPoaError as e:
        return ("err", str(e))


def _take_records(h):
    """Copies an mpoa_records handle into Sequence records and frees it."""
    L = lib()
    out = []
    try:
        for k in range(L.mpoa_records_len(h)):
            ptr = [C.c_void_p() for _ in range(4)]
            ln = [C.c_uint64() for _ in range(4)]
            _check(L.mpoa_record_get(h, k, C.byref(ptr[0]), C.byref(ln[0]), C.byref(ptr[1]),
                                     C.byref(ln[1]), C.byref(ptr[2]), C.byref(ln[2]),
                                     C.byref(ptr[3]), C.byref(ln[3])))
            f = [C.string_at(p, n.value) if n.value else b"" for p, n in zip(ptr, ln)]
            out.append(Sequence(name=f[0].decode(), comment=f[1].decode(), seq=f[2], qual=f[3]))
    finally:
        L.mpoa_records_free(h)
    return out


def parse_sequences(data):
    """`io::parse_sequences` / `io::read_sequences` on bytes (FASTA, FASTQ,
    gzip-compressed either): a list of Sequence records."""
    data = bytes(data)
    h = C.c_void_p()
    _check(lib().mpoa_read_sequences(data, len(data), C.byref(h)))
    return _take_records(h)


def read_sequences(path):
    """`io::read_sequences` (the --input file)."""
    return parse_sequences(Path(path).read_bytes())


def read_gfa(params, data):
    """`io::read_gfa` (the -i graph) on GFA text (bytes or str): (Graph,
    records), one record per path in file order."""
    if isinstance(data, str):
        data = data.encode()
    data = bytes(data)
    cp = CParams(params)
    g, r = C.c_void_p(), C.c_void_p()
    _check(lib().mpoa_read_gfa(cp.ref, data, len(data), C.byref(g), C.byref(r)))
    return Graph(params, _handle=g.value), _take_records(r)


def load_matrix(params, path):
    """`io::load_matrix` (the -m option, src/lib.rs:209-211): reads the
    matrix file at `path` into params.alphabet and params.matrix (header
    order). Raises HipPoaError on malformed text, leaving params unchanged."""
    text = Path(path).read_bytes()
    cap = 256
    alpha = (C.c_uint8 * cap)()
    mat = (C.c_int32 * (cap * cap))()
    n = C.c_uint32(0)
    _check(lib().mpoa_load_matrix(text, len(text), alpha, mat, cap, C.byref(n)))
    k = n.value
    params.alphabet = bytes(alpha[:k])
    params.matrix = list(mat[:k * k])
    return params


def validate(params):
    _check(lib().mpoa_params_validate(CParams(params).ref))


# ---- minimizer seeding (src/minimizer.rs), host side ----

def _sized(call, cap):
    """call(cap) -> (rc, n, result(n)); grows the buffer once if needed."""
    while True:
        rc, n, res = call(cap)
        if rc == BUFFER_TOO_SMALL and n > cap:
            cap = n
            continue
        _check(rc)
        return res(n)


def sketch(seq, k, w):
    """src/minimizer.rs:32-76 -> [Minimizer(value, end, span)]."""
    s = bytes(seq)

    def call(cap):
        v = (C.c_uint64 * max(1, cap))()
        e = (C.c_uint64 * max(1, cap))()
        n = C.c_uint64(0)
        rc = lib().mpoa_sketch(s, len(s), k, w, v, e, cap, C.byref(n))
        return rc, n.value, lambda n: [Minimizer(v[i], e[i], k) for i in range(n)]
    return _sized(call, len(s) + 1)


def _anchor_arr(anchors):
    arr = (_Anchor * max(1, len(anchors)))()
    for i, a in enumerate(anchors):
        arr[i] = _Anchor(a.target_end, a.query_end, a.span, int(bool(a.adaptive_before)))
    return arr


def _anchor_list(arr, n):
    return [Anchor(arr[i].target_end, arr[i].query_end, arr[i].span, bool(arr[i].adaptive_before))
            for i in range(n)]


def _anchor_call(fn, args, cap, tail=None):
    def call(cap):
        out = (_Anchor * max(1, cap))()
        n = C.c_uint64(0)
        extra = (C.byref(tail),) if tail is not None else ()
        rc = fn(*args, out, cap, C.byref(n), *extra)
        return rc, n.value, lambda n: _anchor_list(out, n)
    return _sized(call, cap)


def collect_anchors(params, query, target):
    """src/minimizer.rs:128-176"""
    cp, q, t = CParams(params), bytes(query), bytes(target)
    return _anchor_call(lib().mpoa_collect_anchors, (cp.ref, q, len(q), t, len(t)), 256)


def chain(anchors, k):
    """src/minimizer.rs:199-377"""
    arr = _anchor_arr(anchors)
    return _anchor_call(lib().mpoa_chain, (arr, len(anchors), k), len(anchors) + 1)


def partition_windows(chain_, tlen, qlen, max_drift, max_win):
    """src/minimizer.rs:379-467 -> (boundaries, tail_adaptive)."""
    arr = _anchor_arr(chain_)
    tail = C.c_int(0)
    res = _anchor_call(lib().mpoa_partition_windows, (arr, len(chain_), tlen, qlen, max_drift, max_win),
                       len(chain_) + 1, tail)
    return res, bool(tail.value)


def guide_order(params, seqs):
    """src/minimizer.rs:489-578 (seqs: Sequence records or byte strings)."""
    raw = [bytes(getattr(s, "seq", s)) for s in seqs]
    cp = CParams(params)
    arr = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_uint64 * max(1, len(raw)))(*[len(r) for r in raw])
    out = (C.c_uint64 * max(1, len(raw)))()
    _check(lib().mpoa_guide_order(cp.ref, arr, lens, len(raw), out))
    return list(out)[:len(raw)]


def anchor_windows(params, query, consensus):
    """src/minimizer.rs:475-488 -> (boundaries, tail_adaptive)."""
    cp, q, t = CParams(params), bytes(query), bytes(consensus)
    tail = C.c_int(0)
    res = _anchor_call(lib().mpoa_anchor_windows, (cp.ref, q, len(q), t, len(t)), 256, tail)
    return res, bool(tail.value)


def align_sequences(ctx, params, seqs):
    """src/lib.rs:177-182 with every DP on the device."""
    cp = CParams(params)
    raw = [bytes(s.seq) for s in seqs]
    arr = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_uint64 * max(1, len(raw)))(*[len(r) for r in raw])
    h = C.c_void_p()
    _check(lib().mpoa_align_sequences(ctx._h, cp.ref, arr, lens, len(raw), C.byref(h)))
    return Graph(params, _handle=h)


def extend_graph(ctx, params, graph, seqs, existing):
    cp = CParams(params)
    raw = [bytes(s.seq) for s in seqs]
    arr = (C.c_char_p * max(1, len(raw)))(*raw)
    lens = (C.c_uint64 * max(1, len(raw)))(*[len(r) for r in raw])
    _check(lib().mpoa_extend_graph(ctx._h, cp.ref, graph._h, arr, lens, len(raw), existing))


class BatchResult:
    def __init__(self, handle, njobs):
        self._h = handle
        self.njobs = njobs

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mpoa_batch_free(self._h)
            self._h = None

    def stats(self):
        s = _Stats()
        _check(lib().mpoa_batch_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _Stats._fields_}

    def status(self, job):
        return lib().mpoa_batch_job_status(self._h, job)

    def output(self, job, mode):
        cap = 1 << 16
        while True:
            buf = C.create_string_buffer(cap)
            n = C.c_uint64(0)
            rc = lib().mpoa_batch_job_output(self._h, job, mode, buf, cap, C.byref(n))
            if rc == BUFFER_TOO_SMALL:
                cap = n.value + 1
                continue
            _check(rc)
            return buf.raw[:n.value].decode()


def _text_out(fn, *args):
    cap = 1 << 16
    while True:
        buf = C.create_string_buffer(cap)
        n = C.c_uint64(0)
        rc = fn(*args, buf, cap, C.byref(n))
        if rc == BUFFER_TOO_SMALL:
            cap = n.value + 1
            continue
        _check(rc)
        return buf.raw[:n.value].decode()


class Server:
    """Serving: POA jobs admitted as they arrive (include/minipoa_hip.h).

    submit() copies a job's reads and returns a ticket at once; the job runs
    in the continuous engine alongside whatever else is admitted. A job may
    extend a graph that already holds `existing` sequences (the -i mode of
    src/lib.rs:212-226, which continues with `extend_graph`).

    resident: keep every admitted job's graph and DP descriptors on the
    device between reads, in slots of size-class step tables whose captured
    step graphs advance all admitted jobs one read per replay."""

    def __init__(self, ctx, params, admit_bytes=0, slots=0, render=True, resident=False):
        self._cp = CParams(params)
        self.params = params
        o = _ServerOpts(admit_bytes, slots, 1 if render else 0, 1 if resident else 0)
        h = C.c_void_p()
        _check(lib().mpoa_server_create(ctx._h, self._cp.ref, C.byref(o), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mpoa_server_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, reads, graph=None, existing=0):
        """reads: list of bytes. graph: a Graph holding `existing` sequences to extend."""
        import numpy as np

        lens = [len(r) for r in reads]
        off = np.zeros(len(reads) + 1, dtype=np.uint64)
        off[1:] = np.cumsum(lens, dtype=np.uint64)
        buf = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", dtype=np.uint8)
        t = C.c_uint64(0)
        _check(lib().mpoa_server_submit(self._h, buf.ctypes.data_as(C.c_void_p),
                                        off.ctypes.data_as(U64P), len(reads),
                                        graph._h if graph is not None else None, existing,
                                        C.byref(t)))
        return t.value

    def done(self, ticket):
        rc = lib().mpoa_server_done(self._h, ticket)
        if rc < 0 or rc > 1:
            _check(rc)
        return rc == 1

    def wait(self, ticket):
        _check(lib().mpoa_server_wait(self._h, ticket))

    def status(self, ticket):
        return lib().mpoa_server_job_status(self._h, ticket)

    def output(self, ticket, mode):
        return _text_out(lib().mpoa_server_job_output, self._h, ticket, mode)

    def times(self, ticket):
        t = _JobTimes()
        _check(lib().mpoa_server_job_times(self._h, ticket, C.byref(t)))
        return {f: getattr(t, f) for f, _ in _JobTimes._fields_}

    def graph(self, ticket):
        h = C.c_void_p()
        _check(lib().mpoa_server_job_graph(self._h, ticket, C.byref(h)))
        return Graph(self.params, _handle=h)

    def stats(self):
        s = _Stats()
        _check(lib().mpoa_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _Stats._fields_}

    def now_ms(self):
        return lib().mpoa_server_now_ms(self._h)


def batch_run(ctx, params, job_off, read_off, bases, resident=False):
    """Batched progressive POA over numpy job/read offsets and a base buffer.

    resident: keep every job's graph and DP descriptors on the device between
    reads, with graph fusion and planning on the device and each length
    bucket's step loop a replayed HIP graph (mpoa_batch_run_resident); the
    outputs are the same."""
    import numpy as np

    job_off = np.ascontiguousarray(job_off, dtype=np.uint64)
    read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    cp = CParams(params)
    h = C.c_void_p()
    fn = lib().mpoa_batch_run_resident if resident else lib().mpoa_batch_run
    _check(fn(ctx._h, cp.ref, len(job_off) - 1, job_off.ctypes.data_as(U64P),
              read_off.ctypes.data_as(U64P), bases.ctypes.data_as(C.c_void_p), C.byref(h)))
    return BatchResult(h, len(job_off) - 1)


def resident_check(params, reads, paths, lds_bytes=0):
    """Runs the device-resident step's graph fusion, re-rank and planning code
    on one host thread beside the host graph and planner, applying paths[r]
    (AlignedBase tuples (node, base, aligned), None for absent) after
    planning read r. Returns "" when every graph and plan matched, else the
    first difference."""
    import numpy as np

    cp = CParams(params)
    raw = [bytes(r) for r in reads]
    read_off = np.zeros(len(raw) + 1, dtype=np.uint64)
    read_off[1:] = np.cumsum([len(r) for r in raw])
    blob = b"".join(raw) or b"\0"
    path_off = np.zeros(len(paths) + 1, dtype=np.uint64)
    path_off[1:] = np.cumsum([len(p) for p in paths])
    flat = [e for p in paths for e in p]
    arr = (_Base * max(1, len(flat)))()
    for k, (node, base, aligned) in enumerate(flat):
        arr[k] = _Base(-1 if node is None else node, base, -1 if aligned is None else aligned)
    rep = C.create_string_buffer(512)
    rc = lib().mpoa_resident_check(cp.ref, blob, read_off.ctypes.data_as(U64P), len(raw), arr,
                                   path_off.ctypes.data_as(U64P), lds_bytes, rep, 512)
    if rc == 0:
        return ""
    return rep.value.decode() or last_error()


__all__ = [
    "AlignedBase", "Params", "Sequence", "records", "default_matrix", "HipPoaError", "Context",
    "parse_sequences", "read_sequences", "read_gfa",
    "Graph", "GraphView", "view_from_graph", "poa", "poa_inner", "poa_kernel", "align",
    "align_sequences", "extend_graph", "batch_run", "backend_name", "i32_scores_safe",
    "validate", "result_or_error", "lib", "last_error", "SIGNATURES", "Server", "load_matrix",
    "Minimizer", "Anchor", "sketch", "collect_anchors", "chain", "partition_windows",
    "anchor_windows", "guide_order", "resident_check",
]
