"""Device DP parity: the gfx950 kernels called through the C-ABI against the
oracle on the same graphs and queries -- paths, errors and final scores
bit-exact -- plus the reference's independent optima.

Reference test cases: src/align.rs:1426-1524, 1603-1706, 1802-1871,
tests/regressions.rs, tests/avx512.rs:115-143, tests/io_cli.rs:421-451.
"""
import random

import pytest

import minipoa_hip as mh
from helpers import (Lcg, alignment_score, dag_dp_optimum, dna, pairwise_optimum, validate_gfa,
                     validate_msa)
from minipoa_hip import Params, Sequence, records, synth
from test_oracle import BAND_SETTINGS, SETTINGS_1080, _branched_cases, cases_1080

pytestmark = pytest.mark.gpu

# test-data/HOXD70.mtx (A C G T N), a fixture the reference's differential
# runs use with -O105 -E12 (scripts/differential.py:76-77).
HOXD70 = [91, -114, -31, -123, 0,
          -114, 100, -125, -31, 0,
          -31, -125, 100, -114, 0,
          -123, -31, -144, 91, 0,
          0, 0, 0, 0, 100]


def linear(oracle, params, reference):
    g = oracle.Graph(params)
    g.add_path(0, oracle.align(params, g, reference))
    g.toposort(False)
    return g


def gv(og, params):
    return mh.view_from_graph(og, len(params.alphabet))


def rank_base(og):
    def f(i):
        n = og.node(i)
        return n["rank"], n["base"]
    return f


def test_device_pairwise_optimum(ctx, oracle):
    """src/align.rs:1426-1455 on the device: path == oracle, score == Gotoh."""
    r = Lcg(1923827, 1442695040888963407)
    params = Params(f=0)
    for _ in range(500):
        reference = bytes(b"ACGT"[r.next() % 4] for _ in range(1 + r.next() % 22))
        query = bytes(b"ACGT"[r.next() % 4] for _ in range(r.next() % 23))
        og = linear(oracle, params, reference)
        path = mh.align(ctx, params, gv(og, params), query)
        assert path == oracle.align(params, og, query)
        assert len(path) == len(query) + 1
        assert alignment_score(params, rank_base(og), path, len(reference)) == \
            pairwise_optimum(params, reference, query)


def test_device_zero_gaps_and_ambiguous_bases(ctx, oracle):
    """src/align.rs:1456-1492 on the device."""
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
            og = linear(oracle, params, reference)
            path = mh.align(ctx, params, gv(og, params), query)
            assert path == oracle.align(params, og, query)
            assert alignment_score(params, rank_base(og), path, len(reference)) == \
                pairwise_optimum(params, reference, query)


def test_device_bands_preserve_simple_indels(ctx, oracle):
    """src/align.rs:1493-1524 (static and adaptive bands) on the device."""
    params = Params(b=2, f=40)
    reference = b"ACGTACGATAGCTACGATCGATGCTAGCTAGCTAGCATCGATCG"
    query = b"ACGTACGATAGCTACGATCGATTTTGCTAGCTAGCTAGCATCGATCG"
    og = linear(oracle, params, reference)
    optimum = pairwise_optimum(params, reference, query)
    for adaptive in (False, True):
        params.adaptive = adaptive
        path = mh.align(ctx, params, gv(og, params), query)
        assert path == oracle.align(params, og, query)
        assert alignment_score(params, rank_base(og), path, len(reference)) == optimum


def test_device_kernels_match_oracle_on_branched_graphs_and_bands(ctx, oracle):
    """src/align.rs:1603-1669: the device i32 and i64 kernels equal the
    oracle's i32 and i64 kernels (paths, errors and final score)."""
    count = 0
    for case, params, template, reads in _branched_cases():
        og = linear(oracle, params, template)
        for rid, query in enumerate(reads, start=1):
            view = gv(og, params)
            for f, adaptive, band in BAND_SETTINGS:
                params.f, params.adaptive = f, adaptive
                ref = oracle.result_or_error(oracle.poa_scalar_i64, params, og, 0, 1, query,
                                             adaptive, band)
                ref_score = oracle.last_score() if ref[0] == "ok" else None
                assert oracle.result_or_error(oracle.poa_simd_i32, params, og, 0, 1, query,
                                              adaptive, band) == ref
                for bits in (32, 64):
                    got = mh.result_or_error(mh.poa_kernel, ctx, params, view, 0, 1, query,
                                             adaptive, band, bits, score=True)
                    if ref[0] == "ok":
                        assert got[0] == "ok", (case, rid, f, adaptive, band, bits, got)
                        assert got[1][0] == ref[1], (case, rid, f, adaptive, band, bits)
                        assert got[1][1] == ref_score, (case, rid, f, adaptive, band, bits)
                    else:
                        assert got == ref, (case, rid, f, adaptive, band, bits)
                    count += 1
            params.f = 0
            path = oracle.poa_scalar_i64(params, og, 0, 1, query, False, 100)
            og.add_path(rid, path)
            og.toposort(False)
    assert count == 35 * 4 * 4 * 2


def test_device_tracebacks_across_vector_and_band_edges(ctx, oracle):
    """src/align.rs:1802-1871: 1,080 cases, device i32 == oracle i64."""
    count = 0
    for length, params, reference, alternate, queries in cases_1080():
        og = linear(oracle, params, reference)
        og.add_path(1, oracle.poa_scalar_i64(params, og, 0, 1, alternate, False, 100))
        og.toposort(False)
        view = gv(og, params)
        for query in queries:
            for f, adaptive, band in SETTINGS_1080:
                params.f, params.adaptive = f, adaptive
                ref = oracle.result_or_error(oracle.poa_scalar_i64, params, og, 0, 1, query,
                                             adaptive, band)
                got = mh.result_or_error(mh.poa_kernel, ctx, params, view, 0, 1, query, adaptive,
                                         band, 32)
                assert got == ref, (length, len(query), f, adaptive, band)
                count += 1
    assert count == 1080


def test_device_full_dp_score_is_brute_force_optimum(ctx, oracle):
    """Final M(n-1, m-1) of both device kernels equals the brute-force optimum
    over every path of small branched DAGs."""
    r = Lcg(424242)
    for case in range(40):
        params = Params(f=0)
        if case % 3 == 1:
            params.gap_open, params.gap_ext = -3, -1
        template = bytes(b"ACGT"[r.next() % 4] for _ in range(3 + r.next() % 6))
        og = linear(oracle, params, template)
        for rid in range(1, 3):
            read = bytearray()
            for base in template:
                x = r.next() % 10
                if x == 0:
                    continue
                read.append(b"ACGT"[r.next() % 4] if x == 1 else base)
                if r.next() % 8 == 0:
                    read.append(b"ACGT"[r.next() % 4])
            og.add_path(rid, oracle.poa_scalar_i64(params, og, 0, 1, bytes(read), False, 100))
            og.toposort(False)
        query = bytes(b"ACGT"[r.next() % 4] for _ in range(r.next() % 9))
        nodes = [og.node(i) for i in range(og.n_nodes)]
        best = dag_dp_optimum(params, [n["base"] for n in nodes], [n["outgoing"] for n in nodes],
                              query)
        for bits in (32, 64):
            _, score = mh.poa_kernel(ctx, params, gv(og, params), 0, 1, query, False, 100, bits,
                                     score=True)
            assert score == best, (case, bits)


def test_device_score_guard_routes_extreme_scores_to_i64(ctx, oracle):
    """src/align.rs:1671-1692"""
    params = Params()
    og = linear(oracle, params, b"ACGTACGTACGT")
    view = gv(og, params)
    assert mh.i32_scores_safe(params, view, 0, 1, 12)
    extreme = params.clone(matrix=[2**31 - 1] * 25)
    assert not mh.i32_scores_safe(extreme, view, 0, 1, 12)
    for p in (params, extreme):
        want = oracle.result_or_error(oracle.poa_scalar_i64, p, og, 0, 1, b"ACGTACCTACGT", False,
                                      100)
        assert mh.result_or_error(mh.poa_inner, ctx, p, view, 0, 1, b"ACGTACCTACGT", False,
                                  100) == want
    with pytest.raises(mh.HipPoaError):
        mh.poa_kernel(ctx, extreme, view, 0, 1, b"ACGT", False, 100, 32)


def test_device_anchor_errors_match_reference_messages(ctx, oracle):
    """src/align.rs:596-608 and 1694-1706 (beg > end is an error)."""
    params = Params()
    og = linear(oracle, params, b"ACGTACGTACGTACGT")
    view = gv(og, params)
    for beg, end in [(1, 0), (0, 99), (99, 1)]:
        want = oracle.result_or_error(oracle.poa_inner, params, og, beg, end, b"A", False, 100)
        got = mh.result_or_error(mh.poa_inner, ctx, params, view, beg, end, b"A", False, 100)
        assert got == want and got[0] == "err"


def test_device_narrow_band_failure_retries_full_dp(ctx, oracle):
    """src/align.rs:192-211: band 0 at vector and band edges leaves the band
    (src/align.rs:1848-1857 cases); `poa` then re-runs the exact full matrix.
    The device reproduces the failure and the retried result."""
    retried = 0
    for length, params, reference, alternate, queries in cases_1080():
        og = linear(oracle, params, reference)
        og.add_path(1, oracle.poa_scalar_i64(params, og, 0, 1, alternate, False, 100))
        og.toposort(False)
        view = gv(og, params)
        for query in queries:
            for band in (0, 1):
                params.f, params.adaptive = 40, False
                inner = oracle.result_or_error(oracle.poa_inner, params, og, 0, 1, query, False,
                                               band)
                retried += inner[0] == "err"
                assert mh.result_or_error(mh.poa_inner, ctx, params, view, 0, 1, query, False,
                                          band) == inner
                assert mh.result_or_error(mh.poa, ctx, params, view, 0, 1, query, False, band) \
                    == oracle.result_or_error(oracle.poa, params, og, 0, 1, query, False, band)
    assert retried > 0


def _pipeline_settings():
    hox = Params(gap_open=-105, gap_ext=-12, matrix=list(HOXD70))
    return [
        ("default", Params()),
        ("full", Params(f=0)),
        ("adaptive", Params(adaptive=True)),
        ("scores", Params(match_score=3, mismatch=-5, gap_open=-6, gap_ext=-1).rebuild_matrix()),
        ("hoxd70", hox),
        ("narrow", Params(b=3, f=40, adaptive=True)),
    ]


def test_device_align_sequences_matches_oracle_bytes(ctx, oracle):
    """Whole progressive pipeline (src/lib.rs:177-204) with every DP on the
    device: consensus, MSA and GFA bytes equal the oracle's (-r0/-r1/-r2)."""
    rng = random.Random(9027)
    for name, params in _pipeline_settings():
        for case in range(4):
            template = dna(80 + 60 * case, seed=hash((name, case)) & 0xFFFF)
            reads = []
            for _ in range(3 + case):
                r = bytearray()
                for c in template:
                    x = rng.random()
                    if x < 0.04:
                        continue
                    r.append(rng.choice(b"ACGT") if x < 0.10 else c)
                    if rng.random() < 0.04:
                        r.append(rng.choice(b"ACGT"))
                reads.append(bytes(r) or b"A")
            seqs = records(reads)
            dg = mh.align_sequences(ctx, params, seqs)
            og = oracle.align_sequences(params, seqs)
            assert dg.consensus(params) == og.consensus(params), (name, case)
            assert dg.msa(params, seqs) == og.msa(params, seqs), (name, case)
            assert dg.gfa(params, seqs) == og.gfa(params, seqs), (name, case)


def test_device_pipeline_known_answers(ctx):
    """tests/io_cli.rs:421-451, 542-561, 612-630 and tests/regressions.rs."""
    p = Params()
    reads = [Sequence("a", "", b"ACGT"), Sequence("b", "", b"ACGT")]
    g = mh.align_sequences(ctx, p, reads)
    assert g.consensus(p)[0] == b"ACGT"
    assert g.msa(p, reads) == ">a \nACGT\n>b \nACGT\n"
    assert "P\ta\t1+,2+,3+,4+\t*" in g.gfa(p, reads)
    rna = Params(is_rna=True)
    assert mh.align_sequences(ctx, rna, [Sequence("rna", "", b"acgu")]).consensus(rna)[0] == b"ACGU"
    reads = [Sequence("a", "", b"ACGT"), Sequence("empty", "", b"")]
    assert mh.align_sequences(ctx, p, reads).msa(p, reads) == ">a \nACGT\n>empty \n----\n"
    z = Params(f=0, match_score=5, mismatch=-1, gap_open=0, gap_ext=-1).rebuild_matrix()
    reads = [Sequence(f"r{i}", "", s) for i, s in enumerate([b"GCANCCTTG", b"CTTCA"])]
    rows = [l.replace("-", "") for l in mh.align_sequences(ctx, z, reads).msa(z, reads).splitlines()
            if not l.startswith(">")]
    assert rows == ["GCANCCTTG", "CTTCA"]
    reads = [Sequence(f"r{i}", "", s) for i, s in enumerate([b"ACGT", b"ACCT"])]
    g = mh.align_sequences(ctx, p, reads)
    assert "ACCT" in g.msa(p, reads)
    mh.extend_graph(ctx, p, g, reads + [Sequence("r2", "", b"ACGT")], 2)
    assert g.consensus(p)[0] == b"ACGT"


def test_device_lane_boundary_workflows_preserve_every_read(ctx):
    """tests/avx512.rs:115-143 on the device."""
    for length in [1, 7, 8, 9, 15, 16, 17, 23, 24, 25, 31, 32, 33, 47, 48, 49, 63, 64, 65, 79]:
        reference = dna(length)
        split = length // 2
        insertion = reference[:split] + b"TGCATGCA" + reference[split:]
        deletion = bytes(b for i, b in enumerate(reference) if i % 5 != 2)
        reads = records([reference, insertion, deletion, b""])
        for adaptive in (False, True):
            params = Params(b=1, f=40, adaptive=adaptive)
            g = mh.align_sequences(ctx, params, reads)
            validate_msa(g.msa(params, reads), reads)
            validate_gfa(g.gfa(params, reads), reads)


def test_device_plain_consensus_config_matches_oracle(ctx, oracle):
    """The reference's CPU-runnable configuration: 2,000 bp template, 10
    reads at 10% error (60/20/20), default Params, -r0/-r1/-r2 bytes."""
    reads = synth.c1()
    seqs = records(reads)
    p = Params()
    dg = mh.align_sequences(ctx, p, seqs)
    og = oracle.align_sequences(p, seqs)
    assert dg.consensus(p) == og.consensus(p)
    assert dg.msa(p, seqs) == og.msa(p, seqs)
    assert dg.gfa(p, seqs) == og.gfa(p, seqs)
