"""Seeded alignment on the device: DPs anchored between interior graph nodes,
and the windows between chained minimizer anchors (src/align.rs:99-188,
src/minimizer.rs), through the one-call path (mpoa_align / align_sequences:
all windows of a read in one launch) and the engine (batch and server: the
windows of every ready job share launches). Paths and rendered bytes equal
the oracle's.

Reference cases: src/align.rs:1525-1559, tests/avx512.rs:145-180.
"""
import random

import numpy as np
import pytest

import minipoa_hip as mh
from helpers import Lcg, dna, validate_gfa, validate_msa
from minipoa_hip import Params, records, synth

pytestmark = pytest.mark.gpu


def fnv1a(data, h=14695981039346656037):
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_interior_anchored_poa_matches_oracle(ctx, oracle):
    """poa(beg, end) between interior nodes of a branched graph, the DP every
    seeded window runs: path or error equal to the oracle's, for static,
    adaptive and full bands."""
    rng = random.Random(77)
    params0 = Params()
    tmpl = np.frombuffer(dna(700), dtype=np.uint8) % 4
    reads = [synth.mutate(np.random.Generator(np.random.PCG64(s)), tmpl, 0.08) for s in range(6)]
    og = oracle.align_sequences(params0, records(reads))
    view = mh.view_from_graph(og, len(params0.alphabet))
    cons = og.consensus(params0)
    ids = cons[1]
    assert len(ids) > 600
    n = 0
    for params, adaptive, band in [(Params(), False, 30), (Params(), True, 100), (Params(f=0), False, 100),
                                   (Params(b=2), False, 2)]:
        for _ in range(25):
            i = rng.randrange(0, len(ids) - 50)
            j = rng.randrange(i + 1, min(len(ids), i + 400))
            beg, end = ids[i], ids[j]
            q = bytes(reads[rng.randrange(6)][i:i + rng.randint(0, j - i + 20)])
            want = oracle.result_or_error(oracle.poa, params, og, beg, end, q, adaptive, band)
            got = mh.result_or_error(mh.poa, ctx, params, view, beg, end, q, adaptive, band)
            assert got[0] == want[0], (beg, end, len(q), got, want)
            if want[0] == "ok":
                assert [tuple(x) for x in got[1]] == [tuple(x) for x in want[1]]
                n += 1
    assert n > 60


def test_seeded_windows_on_a_linear_graph(ctx, oracle):
    """src/align.rs:1525-1559: more than two windows; path == oracle's."""
    p = Params(seeding=True, k=7, w=4, tr_dist=0, win_max=80, win_band=20)
    r = Lcg(123981)
    reference = bytes(b"ACGT"[r.next() % 4] for _ in range(2400))
    q = bytearray(reference)
    q[800:800] = b"TGCATGC"
    del q[1600:1605]
    q[2000] = ord("C") if q[2000] == ord("A") else ord("A")
    q = bytes(q)
    og = oracle.Graph(p)
    og.add_path(0, oracle.align(p, og, reference))
    og.toposort(False)
    assert len(mh.anchor_windows(p, q, reference)[0]) > 2
    want = oracle.align(p, og, q)
    got = mh.align(ctx, p, mh.view_from_graph(og, len(p.alphabet)), q)
    assert len(got) == len(q) + 1
    assert [tuple(x) for x in got] == [tuple(x) for x in want]


def test_seeded_workflow_bytes(ctx, oracle):
    """tests/avx512.rs:145-180 (progressive=false) on the device."""
    reference = dna(1409)
    insertion = reference[:417] + dna(25) + reference[417:]
    deletion = reference[:803] + reference[820:]
    reads = records([reference, insertion, deletion])
    p = Params(seeding=True, k=7, w=3, tr_dist=0, win_max=80, win_band=24)
    g = mh.align_sequences(ctx, p, reads)
    og = oracle.align_sequences(p, reads)
    assert g.consensus(p)[0] == og.consensus(p)[0]
    msa, gfa = g.msa(p, reads), g.gfa(p, reads)
    assert msa == og.msa(p, reads)
    assert gfa == og.gfa(p, reads)
    validate_msa(msa, reads)
    validate_gfa(gfa, reads)


SEEDED_SETTINGS = [
    Params(seeding=True),                                              # defaults: k19 w10, 500-base windows
    Params(seeding=True, k=15, w=5, win_band=50, win_max=300),         # many short windows
    Params(seeding=True, k=11, w=3, win_band=20, win_max=120, tr_dist=0, adaptive=True),
    Params(seeding=True, k=15, w=5, win_band=40, win_max=200, f=0),    # full-row windows
]


@pytest.mark.parametrize("si", range(len(SEEDED_SETTINGS)))
def test_seeded_batch_matches_oracle_bytes(ctx, oracle, si):
    """Seeded jobs (and one job of unrelated reads, which has no chain and
    takes the whole-graph fallback) through the batch engine."""
    p = SEEDED_SETTINGS[si]
    jo, ro, b = synth.seeded(seed=5 + si, njobs=5, length=2500, nreads=7, trim=200)
    rng = np.random.Generator(np.random.PCG64(99))
    unrelated = [synth.ACGT[rng.integers(0, 4, size=900)].tobytes() for _ in range(3)]
    per_job = [synth.job_reads(jo, ro, b, j) for j in range(len(jo) - 1)] + [unrelated]
    jo, ro, b = synth._pack(per_job)
    res = mh.batch_run(ctx, p, jo, ro, b)
    for j, reads in enumerate(per_job):
        seqs = records(reads)
        og = oracle.align_sequences(p, seqs)
        assert res.status(j) == 0, mh.last_error()
        assert res.output(j, 0) == ">consensus sequence\n" + og.consensus(p)[0].decode() + "\n"
        assert res.output(j, 1) == og.msa(p, seqs)
        assert res.output(j, 2) == og.gfa(p, seqs)


def test_seeded_server_jobs_match_oracle(ctx, oracle):
    p = Params(seeding=True, k=15, w=5, win_band=60, win_max=300)
    jo, ro, b = synth.seeded(seed=31, njobs=4, length=1800, nreads=6, trim=150)
    per_job = [synth.job_reads(jo, ro, b, j) for j in range(len(jo) - 1)]
    with mh.Server(ctx, p) as srv:
        tickets = [srv.submit(reads) for reads in per_job]
        for t, reads in zip(tickets, per_job):
            srv.wait(t)
            assert srv.status(t) == 0
            seqs = records(reads)
            og = oracle.align_sequences(p, seqs)
            assert srv.output(t, 1) == og.msa(p, seqs)
            assert srv.output(t, 2) == og.gfa(p, seqs)


def test_seeded_long_read_sample_matches_oracle_hash(ctx, oracle):
    """The seeded long-read workload (16.5 kb template, reads at 10% with
    up to 500 bases trimmed at each end), first 12 reads: consensus + MSA
    hash equal to the oracle's."""
    jo, ro, b = synth.seeded(seed=20263, njobs=1, nreads=12)
    p = Params(seeding=True)
    res = mh.batch_run(ctx, p, jo, ro, b)
    cells, want = oracle.batch_run(p, jo, ro, b, 1)
    got = fnv1a((res.output(0, 0) + res.output(0, 1)).encode())
    assert want[0] != 0
    assert got == int(want[0])
    assert res.stats()["cells"] == cells


def guided_reads(seed, n=7):
    """Reads of two diverged templates, interleaved: the -p order groups them."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = [rng.integers(0, 4, size=L).astype(np.uint8) for L in (1800, 1500)]
    return [synth.mutate(rng, t[k % 2], 0.06) for k in range(n)]


@pytest.mark.parametrize("seeding", [False, True])
def test_progressive_guide_order_matches_oracle_bytes(ctx, oracle, seeding):
    """-p (src/lib.rs:189-193, src/minimizer.rs:489-578) through the one-call
    path, the batch engine and the server."""
    p = Params(progressive=True, seeding=seeding, k=15, w=5, win_band=60, win_max=300,
               mm_filter_ratio=0.6)
    per_job = [guided_reads(s) for s in (1, 2, 3)]
    assert any(oracle.guide_order(p, r) != list(range(len(r))) for r in per_job)
    want = []
    for reads in per_job:
        seqs = records(reads)
        og = oracle.align_sequences(p, seqs)
        want.append((">consensus sequence\n" + og.consensus(p)[0].decode() + "\n", og.msa(p, seqs),
                     og.gfa(p, seqs)))
    for reads, w in zip(per_job, want):
        seqs = records(reads)
        g = mh.align_sequences(ctx, p, seqs)
        assert (g.msa(p, seqs), g.gfa(p, seqs)) == w[1:]
    jo, ro, b = synth._pack(per_job)
    res = mh.batch_run(ctx, p, jo, ro, b)
    for j, w in enumerate(want):
        assert res.status(j) == 0, mh.last_error()
        assert (res.output(j, 0), res.output(j, 1), res.output(j, 2)) == w
    with mh.Server(ctx, p) as srv:
        tickets = [srv.submit(r) for r in per_job]
        for t, w in zip(tickets, want):
            srv.wait(t)
            assert srv.status(t) == 0
            assert (srv.output(t, 0), srv.output(t, 1), srv.output(t, 2)) == w


def _oracle_bytes(oracle, p, reads):
    seqs = records(reads)
    og = oracle.align_sequences(p, seqs)
    return (">consensus sequence\n" + og.consensus(p)[0].decode() + "\n", og.msa(p, seqs), og.gfa(p, seqs))


def test_seeded_reads_with_case_ambiguity_and_rna(ctx, oracle):
    """Seeding skips ambiguous runs and treats lower case and U as their
    bases (src/minimizer.rs:28-76); the DP encodes them as the reference
    does. Lower-case/N reads and an RNA job through the batch."""
    rng = np.random.Generator(np.random.PCG64(41))
    t = rng.integers(0, 4, size=2200).astype(np.uint8)
    messy = []
    for k in range(6):
        r = bytearray(synth.mutate(rng, t, 0.05))
        a = int(rng.integers(100, len(r) - 200))
        r[a:a + 25] = b"N" * 25
        for i in rng.choice(len(r), size=len(r) // 10, replace=False):
            r[int(i)] |= 0x20
        messy.append(bytes(r))
    rna = [synth.mutate(rng, t[:1500], 0.05).replace(b"T", b"U") for _ in range(5)]
    for reads, p in ((messy, Params(seeding=True, k=15, w=5, win_band=60, win_max=300)),
                     (rna, Params(seeding=True, k=15, w=5, win_band=60, win_max=300, is_rna=True))):
        assert mh.anchor_windows(p, reads[1], reads[0])[0]
        jo, ro, b = synth._pack([reads])
        res = mh.batch_run(ctx, p, jo, ro, b)
        assert res.status(0) == 0, mh.last_error()
        assert (res.output(0, 0), res.output(0, 1), res.output(0, 2)) == _oracle_bytes(oracle, p, reads)


def test_seeded_windows_split_between_i32_and_i64(ctx, oracle):
    """Scores near the i32 guard (src/align.rs:493-529): each window DP is
    guarded on its own rows and query, so the windows of one step split
    between the i32 and the i64 kernels (separate launches), and the result
    equals the oracle's."""
    # max |score| 110,000: rows + query <= 2^26 / 110,000 - 20 = 590 stays
    # on the i32 kernel; 9 of this job's 69 windows are larger.
    mat = [110000 if a == b else -110000 for a in range(5) for b in range(5)]
    mat[24] = 110000
    p = Params(seeding=True, k=15, w=5, win_band=60, win_max=300, match_score=110000, mismatch=-110000,
               gap_open=-20000, gap_ext=-10000, matrix=mat)
    rng = np.random.Generator(np.random.PCG64(43))
    t = rng.integers(0, 4, size=2600).astype(np.uint8)
    reads = [synth.mutate(rng, t, 0.06) for _ in range(5)]
    jo, ro, b = synth._pack([reads])
    res = mh.batch_run(ctx, p, jo, ro, b)
    st = res.stats()
    assert res.status(0) == 0, mh.last_error()
    assert 0 < st["i64_calls"] < st["dp_calls"]
    assert (res.output(0, 0), res.output(0, 1), res.output(0, 2)) == _oracle_bytes(oracle, p, reads)


def test_seeded_incremental_extension_matches_oracle(ctx, oracle):
    """A served job extends a graph of earlier reads with seeding on
    (src/lib.rs:184-204): equal to aligning all reads at once."""
    p = Params(seeding=True, k=15, w=5, win_band=60, win_max=300)
    rng = np.random.Generator(np.random.PCG64(47))
    t = rng.integers(0, 4, size=2000).astype(np.uint8)
    reads = [synth.mutate(rng, t, 0.08) for _ in range(10)]
    with mh.Server(ctx, p) as srv:
        t0 = srv.submit(reads[:6])
        srv.wait(t0)
        t1 = srv.submit(reads[6:], graph=srv.graph(t0), existing=6)
        srv.wait(t1)
        assert srv.status(t1) == 0
        assert (srv.output(t1, 0), srv.output(t1, 1), srv.output(t1, 2)) == _oracle_bytes(oracle, p, reads)


def test_seeded_edge_reads_match_oracle(ctx, oracle):
    """Seeded jobs with an empty read, a read shorter than k, an all-N read
    (no minimizers: the whole-graph fallback), an exact copy (dense
    anchors), a copy with an inverted segment (the chain breaks: adaptive
    windows) and a truncated copy, through the batch and the one-call path;
    and anchored DPs with an empty query between interior nodes (a window
    whose query gap is zero)."""
    rng = np.random.Generator(np.random.PCG64(53))
    t = synth.ACGT[rng.integers(0, 4, size=1500)].tobytes()
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    inv = t[:500] + t[500:900][::-1].translate(comp) + t[900:]
    reads = [t, b"", b"ACG", b"N" * 300, t, inv, t[100:1400]]
    p = Params(seeding=True, k=11, w=3, win_band=40, win_max=150, tr_dist=0)
    want = _oracle_bytes(oracle, p, reads)
    jo, ro, b = synth._pack([reads])
    res = mh.batch_run(ctx, p, jo, ro, b)
    assert res.status(0) == 0, mh.last_error()
    assert (res.output(0, 0), res.output(0, 1), res.output(0, 2)) == want
    seqs = records(reads)
    g = mh.align_sequences(ctx, p, seqs)
    assert (g.msa(p, seqs), g.gfa(p, seqs)) == want[1:]
    og = oracle.align_sequences(p, seqs)
    view = mh.view_from_graph(og, len(p.alphabet))
    ids = og.consensus(p)[1]
    for i, j in [(0, 1), (10, 11), (100, 140), (600, 1100)]:
        for adaptive, band in [(False, 40), (True, 100)]:
            want1 = oracle.result_or_error(oracle.poa, p, og, ids[i], ids[j], b"", adaptive, band)
            got1 = mh.result_or_error(mh.poa, ctx, p, view, ids[i], ids[j], b"", adaptive, band)
            assert got1[0] == want1[0], (i, j, got1, want1)
            if want1[0] == "ok":
                assert [tuple(x) for x in got1[1]] == [tuple(x) for x in want1[1]]


def test_seeded_window_retries_with_full_matrix(ctx, oracle):
    """With win_band 0 the static band of many windows (0 + m/40) misses the
    path around 40-base indels: each such window's banded DP fails and poa
    retries it with the full matrix (src/align.rs:202-207) -- 77 of 425
    windows on the oracle for these reads. Batch and one-call path equal the
    oracle's bytes."""
    rng = np.random.Generator(np.random.PCG64(61))
    t = rng.integers(0, 4, size=3000).astype(np.uint8)
    reads = []
    for _ in range(6):
        r = synth.mutate(rng, t, 0.03)
        a, b = int(rng.integers(300, 1300)), int(rng.integers(1600, 2600))
        reads.append(r[:a] + r[a + 40:b] + synth.ACGT[rng.integers(0, 4, size=40)].tobytes() + r[b:])
    p = Params(seeding=True, k=15, w=5, win_band=0, win_max=300)
    want = _oracle_bytes(oracle, p, reads)
    jo, ro, b = synth._pack([reads])
    res = mh.batch_run(ctx, p, jo, ro, b)
    assert res.status(0) == 0, mh.last_error()
    assert res.stats()["retries"] > 0
    assert (res.output(0, 0), res.output(0, 1), res.output(0, 2)) == want
    seqs = records(reads)
    g = mh.align_sequences(ctx, p, seqs)
    assert (g.msa(p, seqs), g.gfa(p, seqs)) == want[1:]


def test_seeded_protein_family_takes_the_whole_graph_fallback(ctx, oracle):
    """Seeding with a 20-letter alphabet has no anchors (src/minimizer.rs:
    129-131), so every read aligns to the whole graph with the adaptive band
    (src/align.rs:101-103); bytes equal the oracle's."""
    sp, jo, ro, b = synth.protein_families(nfam=3, lmin=120, lmax=200, members=6)
    p = Params(seeding=True, alphabet=sp.alphabet, matrix=sp.matrix, f=40)
    res = mh.batch_run(ctx, p, jo, ro, b)
    for j in range(len(jo) - 1):
        reads = synth.job_reads(jo, ro, b, j)
        assert res.status(j) == 0, mh.last_error()
        assert (res.output(j, 0), res.output(j, 1), res.output(j, 2)) == _oracle_bytes(oracle, p, reads)
