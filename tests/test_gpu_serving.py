"""Serving on the device: jobs submitted while others run are admitted by
device footprint and share launches; every job's consensus / MSA / GFA bytes
equal the oracle's for the same reads, and equal the job's result when it
runs alone. Incremental jobs extend an existing graph as the -i mode does
(src/lib.rs:212-226 -> extend_graph, src/lib.rs:184-204)."""
import time

import numpy as np
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, records, synth

pytestmark = pytest.mark.gpu


def small_stream(seed, njobs):
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    for _ in range(njobs):
        L = int(rng.integers(60, 900))
        R = int(rng.integers(2, 9))
        tmpl = rng.integers(0, 4, size=L).astype(np.uint8)
        jobs.append([synth.mutate(rng, tmpl, float(rng.uniform(0.05, 0.15))) for _ in range(R)])
    return jobs


def outputs(srv, t):
    assert srv.status(t) == 0, mh.last_error()
    return srv.output(t, 0), srv.output(t, 1), srv.output(t, 2)


def oracle_outputs(oracle, params, reads):
    seqs = records(reads)
    og = oracle.align_sequences(params, seqs)
    return (">consensus sequence\n" + og.consensus(params)[0].decode() + "\n",
            og.msa(params, seqs), og.gfa(params, seqs))


@pytest.mark.parametrize("params", [Params(adaptive=True), Params(), Params(f=0)],
                         ids=["adaptive", "static", "full"])
def test_results_same_alone_in_the_mix_and_as_oracle(ctx, oracle, params):
    jobs = small_stream(31, 14)
    mixed = {}
    with mh.Server(ctx, params) as srv:
        tickets = []
        for k, reads in enumerate(jobs):
            tickets.append(srv.submit(reads))
            if k % 3 == 2:
                time.sleep(0.002)  # later jobs arrive while earlier ones run
        for k, t in enumerate(tickets):
            srv.wait(t)
            mixed[k] = outputs(srv, t)
            tm = srv.times(t)
            assert tm["submit_ms"] <= tm["admit_ms"] <= tm["done_ms"]
    for k, reads in enumerate(jobs):
        assert mixed[k] == oracle_outputs(oracle, params, reads), k
    # Alone: a fresh server running one job at a time.
    with mh.Server(ctx, params) as srv:
        for k in (0, 5, 13):
            t = srv.submit(jobs[k])
            srv.wait(t)
            assert outputs(srv, t) == mixed[k], k


def test_incremental_extension_equals_all_reads_at_once(ctx, oracle):
    params = Params(adaptive=True)
    rng = np.random.Generator(np.random.PCG64(5))
    tmpl = rng.integers(0, 4, size=700).astype(np.uint8)
    reads = [synth.mutate(rng, tmpl, 0.1) for _ in range(14)]
    with mh.Server(ctx, params) as srv:
        t0 = srv.submit(reads[:9])
        srv.wait(t0)
        base = srv.graph(t0)
        # The extension and a fresh job run together.
        t1 = srv.submit(reads[9:], graph=base, existing=9)
        t2 = srv.submit(reads[:5])
        srv.wait(t1)
        srv.wait(t2)
        assert outputs(srv, t1) == oracle_outputs(oracle, params, reads)
        assert outputs(srv, t2) == oracle_outputs(oracle, params, reads[:5])
        # The graph handed in is not modified.
        assert base.gfa(params, records(reads[:9])) == srv.output(t0, 2)


def test_admission_by_footprint_serializes_when_memory_is_short(ctx, oracle):
    """With a footprint bound below any job's estimate, one job runs at a
    time in arrival order; results are unchanged."""
    params = Params(adaptive=True)
    jobs = small_stream(77, 5)
    with mh.Server(ctx, params, admit_bytes=1) as srv:
        tickets = [srv.submit(r) for r in jobs]
        for t in tickets:
            srv.wait(t)
        times = [srv.times(t) for t in tickets]
        for a, b in zip(times, times[1:]):
            assert b["admit_ms"] >= a["done_ms"]
        for k, t in enumerate(tickets):
            assert outputs(srv, t) == oracle_outputs(oracle, params, jobs[k]), k


def test_context_slots_are_exclusive_to_one_engine(ctx):
    params = Params()
    with mh.Server(ctx, params):
        with pytest.raises(mh.HipPoaError):
            mh.Server(ctx, params)
        jo, ro, b = synth.jobs(seed=3, njobs=2, lmin=40, lmax=60, rmin=2, rmax=3)
        with pytest.raises(mh.HipPoaError):
            mh.batch_run(ctx, params, jo, ro, b)
    # Released when the server closes.
    jo, ro, b = synth.jobs(seed=3, njobs=2, lmin=40, lmax=60, rmin=2, rmax=3)
    assert mh.batch_run(ctx, params, jo, ro, b).status(0) == 0


def test_long_queries_stage_codes_in_lds(ctx, oracle):
    """Reads longer than 8,192 bases (query codes staged in LDS beside a
    smaller row ring, wide adaptive rows) match the oracle byte for byte."""
    params = Params(adaptive=True)
    rng = np.random.Generator(np.random.PCG64(12))
    tmpl = rng.integers(0, 4, size=12000).astype(np.uint8)
    reads = [synth.mutate(rng, tmpl, 0.12) for _ in range(4)]
    with mh.Server(ctx, params) as srv:
        t = srv.submit(reads)
        srv.wait(t)
        assert outputs(srv, t) == oracle_outputs(oracle, params, reads)


def test_degenerate_jobs_match_oracle(ctx, oracle):
    """A job with no reads (empty graph), one read (no DP: the empty-graph
    shortcut of src/align.rs:82-95) and an empty read inside a job
    (tests/io_cli.rs:612-630) give the oracle's bytes; they run in the same
    stream as ordinary jobs."""
    params = Params()
    cases = [[], [b"ACGTACGTTGCA"], [b"ACGT", b""], [b"", b"ACGT", b"ACGA"],
             [b"ACGTTGCAACGT", b"ACGTTGCACGT", b"ACGTAGCAACGT"]]
    with mh.Server(ctx, params) as srv:
        tickets = [srv.submit(r) for r in cases]
        for t, reads in zip(tickets, cases):
            srv.wait(t)
            assert srv.status(t) == 0, mh.last_error()
            if not reads:
                assert srv.output(t, 0) == ">consensus sequence\n\n"
                continue
            assert outputs(srv, t) == oracle_outputs(oracle, params, reads), reads


def fnv1a(data, h=14695981039346656037):
    for x in data:
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_protein_families_full_dp_match_oracle(ctx, oracle):
    """Protein POA with a drawn symmetric matrix (diagonal U[4,11],
    off-diagonal U[-4,3]), 20-letter alphabet and full DP (-f0): a sample of
    the 2,000-family workload, consensus FASTA + MSA hashes per family equal
    the oracle's."""
    params, jo, ro, b = synth.protein_families(nfam=24, members=8)
    res = mh.batch_run(ctx, params, jo, ro, b)
    cells, want = oracle.batch_run(params, jo, ro, b, 8)
    got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode()) for j in range(24)],
                   dtype=np.uint64)
    assert res.stats()["cells"] == cells
    assert (got == want).all(), np.nonzero(got != want)
