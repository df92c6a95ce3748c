"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host graph (src/graph.rs semantics) matches the
oracle byte for byte. No device compute here."""
import random
import re
from pathlib import Path

import pytest

import minipoa_hip as mh
from helpers import dna, validate_gfa, validate_msa
from minipoa_hip import AlignedBase, Params, Sequence, records

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "minipoa_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpoa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = mh.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(mh.SIGNATURES), set(syms) ^ set(mh.SIGNATURES)


def test_backend_name_identifies_device_kernel():
    """src/lib.rs:15-37 analogue: the backend string names the kernel width."""
    name = mh.backend_name()
    assert name.startswith("hip-gfx950") and "i32" in name


def test_context_creation_fails_cleanly_without_device():
    try:
        ctx = mh.Context(0)
    except mh.HipPoaError as e:
        assert e.status in (13, 14)
    else:
        ctx.close()


def test_layout_setter_rejects_null_context():
    """mpoa_ctx_set_layout without a context is an argument error."""
    assert mh.lib().mpoa_ctx_set_layout(None, mh.LAYOUT_RECORDS) == mh.INVALID_ARGUMENT


def test_server_entry_points_reject_null_arguments():
    """The serving entry points fail cleanly, with no device and no server."""
    import ctypes as C

    L = mh.lib()
    h = C.c_void_p()
    assert L.mpoa_server_create(None, None, None, C.byref(h)) == 17
    t = C.c_uint64(0)
    assert L.mpoa_server_submit(None, None, None, 0, None, 0, C.byref(t)) == 17
    assert L.mpoa_server_wait(None, 0) == 17
    assert L.mpoa_server_done(None, 0) == 17
    assert L.mpoa_server_job_status(None, 0) == 17
    assert L.mpoa_server_job_times(None, 0, None) == 17
    L.mpoa_server_free(None)
    assert L.mpoa_server_now_ms(None) == 0.0


def test_params_validation_messages():
    """src/lib.rs:149-173"""
    mh.validate(Params())
    cases = [
        (Params(alphabet=b"", matrix=[]), "invalid scoring matrix dimensions"),
        (Params(threads=0), "thread count must be positive"),
        (Params(k=0), "k must be 1..31"),
        (Params(gap_open=1), "gap penalties must be zero or negative"),
        (Params(result=3), "result format and verbosity must be 0..2"),
        (Params(mm_filter_ratio=1.5), "minimizer filter ratio"),
        (Params(win_max=0), "window maximum must be positive"),
    ]
    for p, msg in cases:
        with pytest.raises(mh.HipPoaError, match=re.escape(msg)):
            mh.validate(p)


def test_weights_and_consensus_ties_library_graph():
    """src/graph.rs:344-380 on the library graph."""
    p = Params()
    g = mh.Graph(p)
    B = AlignedBase
    g.add_path(0, [B(0, 4, None), B(None, 0, None), B(None, 1, None)])
    g.add_path(1, [B(0, 4, None), B(2, 0, None), B(None, 2, 3)])
    g.add_path(2, [B(0, 4, None), B(2, 0, None), B(4, 2, None)])
    g.toposort(False)
    assert g.consensus(p)[0] == b"AG"
    assert g.consensus(p)[2] == [3, 2]
    assert g.aligned(3, 2) == 4


def test_rejects_cycles_library_graph():
    """src/graph.rs:381-392"""
    p = Params()
    g = mh.Graph(p)
    a = g.add_node(0)
    b = g.add_node(1)
    g.add_edge(0, a)
    g.add_edge(a, b)
    g.add_edge(b, a)
    g.add_edge(b, 1)
    with pytest.raises(mh.HipPoaError, match="cycle"):
        g.toposort(False)


def test_add_path_rejects_bad_records():
    """src/graph.rs:84-93"""
    p = Params()
    g = mh.Graph(p)
    with pytest.raises(mh.HipPoaError, match="outside alphabet"):
        g.add_path(0, [AlignedBase(0, 9, None)])
    with pytest.raises(mh.HipPoaError, match="unknown node"):
        g.add_path(0, [AlignedBase(7, 0, None)])


def test_graph_node_and_edge_exports_check_arguments():
    """The reference panics (Rust bounds checks, src/graph.rs:42-80) on a
    base outside the alphabet or an unknown node id; the C ABI refuses them
    and leaves the graph unchanged."""
    import ctypes as C
    L = mh.lib()
    p = Params()
    g = mh.Graph(p)
    alpha = len(p.alphabet)
    assert L.mpoa_graph_add_node(g._h, alpha) == (1 << 64) - 1
    assert "alphabet" in mh.last_error()
    assert L.mpoa_graph_add_node(g._h, (1 << 32) + 1) == (1 << 64) - 1
    assert L.mpoa_graph_add_node(None, 0) == (1 << 64) - 1
    with pytest.raises(mh.HipPoaError):
        g.add_node(alpha)
    assert g.n_nodes == 2
    a = g.add_node(0)
    assert a == 2
    assert L.mpoa_graph_add_edge(g._h, 0, 3) == 17
    assert L.mpoa_graph_add_edge(g._h, (1 << 32) + 2, a) == 17
    assert L.mpoa_graph_add_edge(None, 0, 1) == 17
    with pytest.raises(mh.HipPoaError, match="unknown node"):
        g.add_edge(a, 99)
    g.add_edge(0, a)
    g.add_edge(a, 1)
    assert g.n_nodes == 3
    assert g.node(a)["incoming"] == [0] and g.node(a)["outgoing"] == [1]


def _same_graph(a, b, alpha):
    assert a.n_nodes == b.n_nodes
    assert a.rank == b.rank
    for i in range(a.n_nodes):
        na, nb = a.node(i), b.node(i)
        for key in ("rank", "parent", "base", "incoming", "in_weights", "outgoing", "out_weights"):
            assert na[key] == nb[key], (i, key)
        for base in range(alpha):
            assert a.aligned(i, base) == b.aligned(i, base)


def test_library_graph_matches_oracle_fusion_and_rendering(oracle):
    """The library graph fed the oracle's alignment paths reproduces the
    oracle's nodes, edges, ranks, consensus, MSA and GFA exactly."""
    rng = random.Random(20261)
    for case in range(12):
        p = Params(f=0 if case % 3 == 0 else 40, adaptive=case % 2 == 1, b=4 + case)
        template = dna(30 + 7 * case, seed=1000 + case)
        reads = [template]
        for _ in range(5):
            r = bytearray()
            for c in template:
                x = rng.random()
                if x < 0.04:
                    continue
                r.append(rng.choice(b"ACGT") if x < 0.10 else c)
                if rng.random() < 0.04:
                    r.append(rng.choice(b"ACGT"))
            reads.append(bytes(r))
        seqs = records(reads)
        og = oracle.Graph(p)
        lg = mh.Graph(p)
        for sid, s in enumerate(seqs):
            path = oracle.align(p, og, s.seq)
            og.add_path(sid, path)
            og.toposort(False)
            lg.add_path(sid, path)
            lg.toposort(False)
        _same_graph(og, lg, 5)
        assert lg.consensus(p) == og.consensus(p)
        assert lg.msa(p, seqs) == og.msa(p, seqs)
        assert lg.gfa(p, seqs) == og.gfa(p, seqs)
        # rendering the MSA restores ranks (tests/regressions.rs:15-34)
        _same_graph(og, lg, 5)
        validate_msa(lg.msa(p, seqs), seqs)
        validate_gfa(lg.gfa(p, seqs), seqs)


def test_library_graph_rna_decoding(oracle):
    """tests/io_cli.rs:542-561: RNA input renders U."""
    p = Params(is_rna=True)
    og = oracle.Graph(p)
    lg = mh.Graph(p)
    path = oracle.align(p, og, b"acgu")
    og.add_path(0, path)
    lg.add_path(0, path)
    og.toposort(False)
    lg.toposort(False)
    assert lg.consensus(p)[0] == og.consensus(p)[0] == b"ACGU"


def test_library_ranks_match_oracle_on_a_large_graph(oracle):
    """toposort(false) (src/graph.rs:124-178) on a branchy graph of 12 reads
    of 2.5 kb at 10% error: the library's ranks equal the oracle's after
    every read."""
    import numpy as np

    from minipoa_hip import synth
    p = Params(adaptive=True)
    rng = np.random.Generator(np.random.PCG64(77))
    t = rng.integers(0, 4, size=2500).astype(np.uint8)
    og, lg = oracle.Graph(p), mh.Graph(p)
    for sid in range(12):
        r = synth.mutate(rng, t, 0.10)
        path = oracle.align(p, og, r)
        og.add_path(sid, path)
        og.toposort(False)
        lg.add_path(sid, path)
        lg.toposort(False)
        assert lg.rank == og.rank, sid
    assert lg.n_nodes > 4000


def test_seeding_entry_points_check_arguments():
    """The seeding exports reject NULL buffers with a status, report the full
    count when the buffer is too small, and accept empty inputs."""
    import ctypes as C
    L = mh.lib()
    n = C.c_uint64(0)
    u64 = C.POINTER(C.c_uint64)
    # NULL sequence with a length; NULL outputs with a capacity.
    assert L.mpoa_sketch(None, 5, 3, 2, None, None, 0, C.byref(n)) == 17
    assert L.mpoa_sketch(b"ACGTACGT", 8, 3, 1, None, None, 4, C.byref(n)) == 17
    # Too small: status 16 and the count needed.
    v = (C.c_uint64 * 1)()
    e = (C.c_uint64 * 1)()
    assert L.mpoa_sketch(b"ACGTACGTAC", 10, 3, 1, v, e, 1, C.byref(n)) == 16
    assert n.value == 8
    # Empty input: no minimizers.
    assert L.mpoa_sketch(b"", 0, 3, 2, v, e, 1, C.byref(n)) == 0 and n.value == 0
    cp = mh.CParams(Params())
    assert L.mpoa_guide_order(cp.ref, None, None, 3, None) == 17
    assert L.mpoa_guide_order(cp.ref, None, None, 0, None) == 0
    assert L.mpoa_chain(None, 2, 19, None, 0, C.byref(n)) == 17
    assert L.mpoa_anchor_windows(None, b"A", 1, b"A", 1, None, 0, C.byref(n), None) == 17
