"""Pins the oracle (the C restatement) against the reference's own test cases,
known answers and independent mathematics. CPU only.

Reference tests restated here: src/align.rs:1426-1706, 1707-1871 (the
avx512 tracebacks run against the 8-lane restatement), src/graph.rs:344-392,
tests/regressions.rs, tests/avx512.rs:115-143, tests/io_cli.rs:421-451.
"""
import pytest

from helpers import (Lcg, alignment_score, dag_dp_optimum, dna, pairwise_optimum,
                     validate_gfa, validate_msa)
from minipoa_hip.params import AlignedBase, Params, Sequence, records

NEG32 = -(1 << 29)


def linear(o, params, reference):
    """src/align.rs:1352-1358"""
    g = o.Graph(params)
    path = o.align(params, g, reference)
    g.add_path(0, path)
    g.toposort(False)
    return g


def rank_base(g):
    cache = {}

    def f(i):
        if i not in cache:
            n = g.node(i)
            cache[i] = (n["rank"], n["base"])
        return cache[i]

    return f


def test_affine_alignment_matches_independent_pairwise_optimum(oracle):
    """src/align.rs:1426-1455"""
    r = Lcg(1923827, 1442695040888963407)
    params = Params(f=0)
    for _ in range(500):
        reference = bytes(b"ACGT"[r.next() % 4] for _ in range(1 + r.next() % 22))
        query = bytes(b"ACGT"[r.next() % 4] for _ in range(r.next() % 23))
        g = linear(oracle, params, reference)
        path = oracle.align(params, g, query)
        assert len(path) == len(query) + 1
        assert alignment_score(params, rank_base(g), path, len(reference)) == \
            pairwise_optimum(params, reference, query), (reference, query)


def test_affine_alignment_handles_zero_gaps_and_ambiguous_bases(oracle):
    """src/align.rs:1456-1492"""
    r = Lcg(912873)
    for matching, mismatch, opening, extension in [(2, -4, -4, -2), (5, -1, 0, -1),
                                                   (0, -2, -2, 0), (3, 0, 0, 0)]:
        params = Params(f=0, gap_open=opening, gap_ext=extension)
        for a in range(4):
            for b in range(4):
                params.matrix[a * 5 + b] = matching if a == b else mismatch
        params.matrix[24] = matching
        for _ in range(150):
            reference = bytes(b"ACGTN"[r.next() % 5] for _ in range(1 + r.next() % 20))
            query = bytes(b"ACGTN"[r.next() % 5] for _ in range(r.next() % 21))
            g = linear(oracle, params, reference)
            path = oracle.align(params, g, query)
            assert len(path) == len(query) + 1
            assert alignment_score(params, rank_base(g), path, len(reference)) == \
                pairwise_optimum(params, reference, query)


def test_static_and_adaptive_bands_preserve_simple_indels(oracle):
    """src/align.rs:1493-1524"""
    params = Params(b=2, f=40)
    reference = b"ACGTACGATAGCTACGATCGATGCTAGCTAGCTAGCATCGATCG"
    query = b"ACGTACGATAGCTACGATCGATTTTGCTAGCTAGCTAGCATCGATCG"
    g = linear(oracle, params, reference)
    full = params.clone(f=0)
    optimum = pairwise_optimum(params, reference, query)
    for adaptive in (False, True):
        params.adaptive = adaptive
        path = oracle.align(params, g, query)
        assert len(path) == len(query) + 1
        assert alignment_score(params, rank_base(g), path, len(reference)) == optimum
    assert alignment_score(params, rank_base(g), oracle.align(full, g, query),
                           len(reference)) == optimum


def test_simd_prefix_scan_matches_serial_affine_transitions(oracle):
    """src/align.rs:1568-1601: the 3-stage shift scan equals the serial
    recurrence I <- max(I + e, md + o)."""
    r = Lcg(8917321)
    for opening, extension in [(0, 0), (0, -2), (-4, 0), (-6, -2), (-117, -12)]:
        for case in range(500):
            md = [NEG32 if r.i31() % 11 == 0 else r.i31() % 201 - 100 for _ in range(8)]
            carry = NEG32 if case % 9 == 0 else r.i31() % 201 - 100
            actual = oracle.insertion_prefix(md, carry, opening, extension)
            value, expected = carry, []
            for x in md:
                value = max(value + extension, x + opening)
                expected.append(value)
            assert actual == expected


def test_prefix_scan_cross_block_carries_and_unreachable(oracle):
    """src/align.rs:1709-1768 (24-column chains through three 8-lane blocks)."""
    r = Lcg(9716271)
    for opening, extension in [(0, 0), (0, -2), (-4, 0), (-6, -2), (-117, -12),
                               (-3_000_000, -3_000_000)]:
        for case in range(400):
            md = [NEG32 if case % 13 == 0 or r.i31() % 7 == 0 else r.i31() % 2001 - 1000
                  for _ in range(24)]
            carry = NEG32 - 1000 if case % 11 == 0 else r.i31() % 2001 - 1000
            actual, c = [], carry
            for blk in range(3):
                out = oracle.insertion_prefix(md[8 * blk:8 * blk + 8], c, opening, extension)
                actual += out
                c = out[7]
            score, expected = carry, []
            for x in md:
                score = max(score + extension, x + opening)
                expected.append(score)
            assert actual == expected


def _branched_cases():
    """src/align.rs:1603-1669 case generator (graphs grown read by read)."""
    r = Lcg(7512129)
    for case in range(35):
        params = Params()
        if case % 4 == 1:
            params.gap_open, params.gap_ext = 0, -1
        elif case % 4 == 2:
            params.gap_ext = 0
        elif case % 4 == 3:
            params.gap_open, params.gap_ext = 0, 0
        if case % 5 == 0:
            params.matrix[24] = 0
        template = bytes(b"ACGTN"[r.next() % 5] for _ in range(40 + r.next() % 80))
        reads = []
        for _rid in range(1, 5):
            query = bytearray()
            for base in template:
                if r.next() % 20 == 0:
                    continue
                query.append(b"ACGTN"[r.next() % 5] if r.next() % 12 == 0 else base)
                if r.next() % 20 == 0:
                    query.append(b"ACGT"[r.next() % 4])
            reads.append(bytes(query))
        yield case, params, template, reads


BAND_SETTINGS = [(0, False, 0), (40, False, 2), (40, True, 3), (40, False, 100)]


def test_simd_matches_i64_traceback_on_branched_graphs_and_bands(oracle):
    """src/align.rs:1603-1669: i32 result == i64 result (paths and errors)."""
    for case, params, template, reads in _branched_cases():
        g = linear(oracle, params, template)
        for rid, query in enumerate(reads, start=1):
            for f, adaptive, band in BAND_SETTINGS:
                params.f, params.adaptive = f, adaptive
                fast = oracle.result_or_error(oracle.poa_simd_i32, params, g, 0, 1, query,
                                              adaptive, band)
                scalar = oracle.result_or_error(oracle.poa_scalar_i64, params, g, 0, 1, query,
                                                adaptive, band)
                assert fast == scalar, (case, rid, f, adaptive, band)
            params.f = 0
            path = oracle.poa_scalar_i64(params, g, 0, 1, query, False, 100)
            g.add_path(rid, path)
            g.toposort(False)


def test_score_guard_dispatches_normal_work_to_simd_and_extreme_work_to_i64(oracle):
    """src/align.rs:1671-1692"""
    params = Params()
    g = linear(oracle, params, b"ACGTACGTACGT")
    oracle.arena_clear()
    assert oracle.i32_scores_safe(params, g, 0, 1, 12)
    oracle.poa_inner(params, g, 0, 1, b"ACGTACCTACGT", False, 100)
    assert oracle.arena_capacity() > 0
    extreme = params.clone(matrix=[2**31 - 1] * 25)
    assert not oracle.i32_scores_safe(extreme, g, 0, 1, 12)
    oracle.arena_clear()
    a = oracle.result_or_error(oracle.poa_inner, extreme, g, 0, 1, b"ACGTACCTACGT", False, 100)
    b = oracle.result_or_error(oracle.poa_scalar_i64, extreme, g, 0, 1, b"ACGTACCTACGT",
                               False, 100)
    assert a == b
    assert oracle.arena_capacity() == 0
    forced = params.clone(scalar_dp=True)
    assert not oracle.i32_scores_safe(forced, g, 0, 1, 12)


def test_simd_arena_reuses_storage_and_error_anchor_order(oracle):
    """src/align.rs:1694-1706 (capacity retained; beg > end is an error)."""
    params = Params()
    g = linear(oracle, params, b"ACGTACGTACGTACGT")
    oracle.poa_simd_i32(params, g, 0, 1, b"ACGTACCTACGTACGT", False, 100)
    before = oracle.arena_capacity()
    oracle.poa_simd_i32(params, g, 0, 1, b"ACGT", False, 100)
    assert oracle.arena_capacity() == before
    with pytest.raises(oracle.OracleError, match="topologically ordered"):
        oracle.poa_simd_i32(params, g, 1, 0, b"A", False, 100)
    assert oracle.arena_capacity() == before


LENGTHS_1080 = [1, 6, 7, 8, 9, 14, 15, 16, 17, 22, 23, 24, 25, 30, 31, 32, 33, 47, 48, 49,
                63, 64, 65, 79, 80, 81, 95]
SETTINGS_1080 = [(0, False, 0), (40, False, 0), (40, False, 1), (40, False, 7),
                 (40, False, 8), (40, False, 15), (40, True, 1), (40, True, 8)]


def cases_1080(oracle_or_none=None):
    """src/align.rs:1802-1871 case generator: (params, graph-builder, query, f,
    adaptive, band) over vector and band edges. Yields exactly 1,080 cases."""
    r = Lcg(1923801)
    template = bytes(b"ACGTN"[r.next() % 5] for _ in range(95))
    for length in LENGTHS_1080:
        params = Params()
        if length % 3 == 0:
            params.gap_open = 0
        if length % 5 == 0:
            params.gap_ext = 0
        reference = template[:length]
        alternate = bytearray(reference)
        if length > 8:
            alternate[7] = ord("C") if alternate[7] == ord("A") else ord("A")
        alternate = bytes(alternate)
        insertion = reference[:length // 2] + b"ACGTACGTA" + reference[length // 2:]
        deletion = bytes(b for i, b in enumerate(reference) if i % 13 != 7)
        queries = [b"", reference, alternate, insertion, deletion]
        yield length, params, reference, alternate, queries


def test_tracebacks_match_scalar_across_vector_and_band_edges(oracle):
    """src/align.rs:1802-1871 -- 1,080 i32 vs i64 comparisons."""
    count = 0
    for length, params, reference, alternate, queries in cases_1080():
        g = linear(oracle, params, reference)
        alt_path = oracle.poa_scalar_i64(params, g, 0, 1, alternate, False, 100)
        g.add_path(1, alt_path)
        g.toposort(False)
        for query in queries:
            for f, adaptive, band in SETTINGS_1080:
                params.f, params.adaptive = f, adaptive
                a = oracle.result_or_error(oracle.poa_simd_i32, params, g, 0, 1, query,
                                           adaptive, band)
                b = oracle.result_or_error(oracle.poa_scalar_i64, params, g, 0, 1, query,
                                           adaptive, band)
                assert a == b, (length, len(query), f, adaptive, band)
                count += 1
    assert count == 1080


def test_fill_score_equals_brute_force_dag_optimum(oracle):
    """Full-DP score on tiny branched DAGs equals the best global affine score
    over every source->sink path (brute force), for both kernels."""
    r = Lcg(424242)
    checked = 0
    for case in range(60):
        params = Params(f=0)
        if case % 3 == 1:
            params.gap_open, params.gap_ext = -3, -1
        elif case % 3 == 2:
            params.gap_open, params.gap_ext = 0, -2
        template = bytes(b"ACGT"[r.next() % 4] for _ in range(3 + r.next() % 6))
        g = linear(oracle, params, template)
        for rid in range(1, 3):
            read = bytearray()
            for base in template:
                x = r.next() % 10
                if x == 0:
                    continue
                read.append(b"ACGT"[r.next() % 4] if x == 1 else base)
                if r.next() % 8 == 0:
                    read.append(b"ACGT"[r.next() % 4])
            g.add_path(rid, oracle.poa_scalar_i64(params, g, 0, 1, bytes(read), False, 100))
            g.toposort(False)
        query = bytes(b"ACGT"[r.next() % 4] for _ in range(r.next() % 9))
        nn = g.n_nodes
        nodes = [g.node(i) for i in range(nn)]
        succ = [n["outgoing"] for n in nodes]
        bases = [n["base"] for n in nodes]
        best = dag_dp_optimum(params, bases, succ, query)
        for fn in (oracle.poa_scalar_i64, oracle.poa_simd_i32):
            fn(params, g, 0, 1, query, False, 100)
            assert oracle.last_score() == best, (case, template, query)
            checked += 1
    assert checked == 120


def test_weights_and_consensus_ties(oracle):
    """src/graph.rs:344-380 known answer."""
    p = Params()
    g = oracle.Graph(p)
    B = AlignedBase
    g.add_path(0, [B(0, 4, None), B(None, 0, None), B(None, 1, None)])
    g.add_path(1, [B(0, 4, None), B(2, 0, None), B(None, 2, 3)])
    g.add_path(2, [B(0, 4, None), B(2, 0, None), B(4, 2, None)])
    g.toposort(False)
    assert g.consensus(p)[0] == b"AG"
    assert g.consensus(p)[2] == [3, 2]
    assert g.node(2)["coverage"] == 3
    assert g.aligned(3, 2) == 4


def test_rejects_cycles(oracle):
    """src/graph.rs:381-392"""
    p = Params()
    g = oracle.Graph(p)
    a = g.add_node(0)
    b = g.add_node(1)
    g.add_edge(0, a)
    g.add_edge(a, b)
    g.add_edge(b, a)
    g.add_edge(b, 1)
    with pytest.raises(oracle.OracleError, match="cycle"):
        g.toposort(False)


def test_pipeline_known_answers(oracle):
    """tests/io_cli.rs:421-451: `>a ACGT >b ACGT` in -r0/-r1/-r2."""
    p = Params()
    reads = [Sequence("a", "", b"ACGT"), Sequence("b", "", b"ACGT")]
    g = oracle.align_sequences(p, reads)
    assert b">consensus sequence\n" + g.consensus(p)[0] + b"\n" == b">consensus sequence\nACGT\n"
    assert g.msa(p, reads) == ">a \nACGT\n>b \nACGT\n"
    gfa = g.gfa(p, reads)
    assert gfa.startswith("H\tVN:Z:1.0\n")
    assert "P\ta\t1+,2+,3+,4+\t*" in gfa


def test_rna_and_empty_record_known_answers(oracle):
    """tests/io_cli.rs:542-561 (RNA acgu -> ACGU) and 612-630 (empty record)."""
    p = Params(is_rna=True)
    g = oracle.align_sequences(p, [Sequence("rna", "", b"acgu")])
    assert g.consensus(p)[0] == b"ACGU"
    p = Params()
    reads = [Sequence("a", "", b"ACGT"), Sequence("empty", "", b"")]
    g = oracle.align_sequences(p, reads)
    assert g.msa(p, reads) == ">a \nACGT\n>empty \n----\n"


def test_rendering_msa_keeps_graph_ready_for_further_alignment(oracle):
    """tests/regressions.rs:15-34"""
    p = Params()
    reads = [Sequence(f"r{i}", "", s) for i, s in enumerate([b"ACGT", b"ACCT"])]
    g = oracle.align_sequences(p, reads)
    order = g.rank
    ranks = [g.node(i)["rank"] for i in range(g.n_nodes)]
    assert "ACCT" in g.msa(p, reads)
    assert g.rank == order
    assert [g.node(i)["rank"] for i in range(g.n_nodes)] == ranks
    path = oracle.align(p, g, b"ACGT")
    g.add_path(2, path)
    g.toposort(False)
    assert g.consensus(p)[0] == b"ACGT"


def test_zero_gap_open_traceback_regression(oracle):
    """tests/regressions.rs:36-55 (-f0 -M5 -X-1 -O0 -E-1 after CLI normalisation)."""
    p = Params(f=0, match_score=5, mismatch=-1, gap_open=0, gap_ext=-1).rebuild_matrix()
    reads = [Sequence(f"r{i}", "", s) for i, s in enumerate([b"GCANCCTTG", b"CTTCA"])]
    g = oracle.align_sequences(p, reads)
    rows = [l.replace("-", "") for l in g.msa(p, reads).splitlines() if not l.startswith(">")]
    assert rows == ["GCANCCTTG", "CTTCA"]


def test_invalid_library_matrix_is_rejected_before_encoding(oracle):
    """tests/regressions.rs:57-67"""
    p = Params(alphabet=b"", matrix=[])
    with pytest.raises(oracle.OracleError, match="matrix"):
        oracle.align_sequences(p, [Sequence("r0", "", b"ACGT")])


def test_lane_boundary_workflows_preserve_every_read(oracle):
    """tests/avx512.rs:115-143"""
    for length in [1, 7, 8, 9, 15, 16, 17, 23, 24, 25, 31, 32, 33, 47, 48, 49, 63, 64, 65, 79]:
        reference = dna(length)
        split = length // 2
        insertion = reference[:split] + b"TGCATGCA" + reference[split:]
        deletion = bytes(b for i, b in enumerate(reference) if i % 5 != 2)
        reads = records([reference, insertion, deletion, b""])
        for adaptive in (False, True):
            params = Params(b=1, f=40, adaptive=adaptive)
            g = oracle.align_sequences(params, reads)
            validate_msa(g.msa(params, reads), reads)
            validate_gfa(g.gfa(params, reads), reads)
