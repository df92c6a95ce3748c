"""Serving on device-resident graphs (mpoa_server_opts.resident): jobs that
arrive while others run are admitted into slots of size-class step tables
as slots and heap room free up, keep their graphs and DP descriptors on the
device between reads, and advance one read per replay of their table's
captured step graph. Every job's bytes must equal the oracle's, and equal
the job's result alone; jobs that leave the device (score guard, arena,
heap) finish on the host engine with the same bytes."""
import os
import time

import numpy as np
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, records, synth

pytestmark = pytest.mark.gpu


def stream(seed, njobs, lmin=60, lmax=900, rmin=2, rmax=9):
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    for _ in range(njobs):
        L = int(rng.integers(lmin, lmax))
        R = int(rng.integers(rmin, rmax))
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


@pytest.fixture
def env():
    saved = {}

    def set_(k, v):
        saved.setdefault(k, os.environ.get(k))
        os.environ[k] = v

    yield set_
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("params", [Params(adaptive=True), Params(), Params(f=0)],
                         ids=["adaptive", "static", "full"])
def test_resident_serving_same_alone_in_the_mix_and_as_oracle(ctx, oracle, params):
    jobs = stream(31, 14)
    mixed = {}
    with mh.Server(ctx, params, resident=True) as srv:
        tickets = []
        for k, reads in enumerate(jobs):
            tickets.append(srv.submit(reads))
            if k % 3 == 2:
                time.sleep(0.003)  # later jobs arrive while earlier ones run
        for k, t in enumerate(tickets):
            srv.wait(t)
            mixed[k] = outputs(srv, t)
            tm = srv.times(t)
            assert tm["submit_ms"] <= tm["admit_ms"] <= tm["done_ms"]
        st = srv.stats()
        assert st["resident_jobs"] == len(jobs) and st["step_replays"] > 0
    for k, reads in enumerate(jobs):
        assert mixed[k] == oracle_outputs(oracle, params, reads), k
    with mh.Server(ctx, params, resident=True) as srv:
        for k in (0, 5, 13):
            t = srv.submit(jobs[k])
            srv.wait(t)
            assert outputs(srv, t) == mixed[k], k


def test_resident_serving_size_classes(ctx, oracle):
    """Templates of 300, 3,000 and 9,000 bases go to three size classes'
    tables and run side by side."""
    params = Params(adaptive=True)
    jobs = stream(41, 2, 250, 350, 3, 5) + stream(42, 2, 2900, 3100, 3, 4) + stream(43, 1, 8900, 9100, 3, 4)
    with mh.Server(ctx, params, resident=True) as srv:
        tickets = [srv.submit(r) for r in jobs]
        for t, reads in zip(tickets, jobs):
            srv.wait(t)
            assert outputs(srv, t) == oracle_outputs(oracle, params, reads)
        assert srv.stats()["resident_jobs"] == len(jobs)


def test_resident_serving_incremental_extension(ctx, oracle):
    """A graph built by one job, extended by a later job (the -i mode,
    src/lib.rs:212-226) while a fresh job runs: the device takes the graph
    with its list orders and paths, and the bytes equal all reads at once."""
    params = Params(adaptive=True)
    rng = np.random.Generator(np.random.PCG64(5))
    tmpl = rng.integers(0, 4, size=700).astype(np.uint8)
    reads = [synth.mutate(rng, tmpl, 0.1) for _ in range(14)]
    with mh.Server(ctx, params, resident=True) as srv:
        t0 = srv.submit(reads[:9])
        srv.wait(t0)
        base = srv.graph(t0)
        t1 = srv.submit(reads[9:], graph=base, existing=9)
        t2 = srv.submit(reads[:5])
        srv.wait(t1)
        srv.wait(t2)
        assert outputs(srv, t1) == oracle_outputs(oracle, params, reads)
        assert outputs(srv, t2) == oracle_outputs(oracle, params, reads[:5])
        assert base.gfa(params, records(reads[:9])) == srv.output(t0, 2)
        assert srv.stats()["resident_jobs"] == 3


def test_resident_serving_degenerate_jobs(ctx, oracle):
    params = Params()
    cases = [[], [b"ACGTACGTTGCA"], [b"ACGT", b""], [b"", b"ACGT", b"ACGA"],
             [b"ACGTTGCAACGT", b"ACGTTGCACGT", b"ACGTAGCAACGT"]]
    with mh.Server(ctx, params, resident=True) as srv:
        tickets = [srv.submit(r) for r in cases]
        for t, reads in zip(tickets, cases):
            srv.wait(t)
            assert srv.status(t) == 0, mh.last_error()
            if not reads:
                assert srv.output(t, 0) == ">consensus sequence\n\n"
                continue
            assert outputs(srv, t) == oracle_outputs(oracle, params, reads), reads


def test_resident_serving_jobs_leave_the_device(ctx, oracle, env):
    """A 1 MiB step arena: DPs that outgrow it leave the device mid-job and
    finish on the host engine; a 1 MiB heap: jobs larger than it never enter.
    Bytes equal the oracle's either way."""
    params = Params(adaptive=True)
    jobs = stream(51, 6, 200, 900, 3, 7)
    env("MPOA_RES_ARENA_MB", "1")
    with mh.Server(ctx, params, resident=True) as srv:
        tickets = [srv.submit(r) for r in jobs]
        for t, reads in zip(tickets, jobs):
            srv.wait(t)
            assert outputs(srv, t) == oracle_outputs(oracle, params, reads)
        assert srv.stats()["evicted_jobs"] > 0
    os.environ.pop("MPOA_RES_ARENA_MB")
    env("MPOA_RES_HEAP_MB", "1")
    jobs = stream(52, 4, 1500, 1900, 5, 7)  # regions of several MB each
    with mh.Server(ctx, params, resident=True) as srv:
        tickets = [srv.submit(r) for r in jobs]
        for t, reads in zip(tickets, jobs):
            srv.wait(t)
            assert outputs(srv, t) == oracle_outputs(oracle, params, reads)
        assert srv.stats()["resident_jobs"] == 0


def test_resident_serving_score_guard_hands_jobs_to_the_host(ctx, oracle):
    p = Params(match_score=1 << 16, mismatch=-(1 << 16), gap_open=-(1 << 16), gap_ext=-(1 << 15)).rebuild_matrix()
    jobs = stream(61, 5, 300, 700, 3, 7)
    with mh.Server(ctx, p, resident=True) as srv:
        tickets = [srv.submit(r) for r in jobs]
        for t, reads in zip(tickets, jobs):
            srv.wait(t)
            assert outputs(srv, t) == oracle_outputs(oracle, p, reads)
        st = srv.stats()
        assert st["evicted_jobs"] > 0 and st["i64_calls"] > 0
