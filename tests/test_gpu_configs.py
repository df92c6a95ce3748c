"""The benchmark configurations at their stated sizes on the device, byte for
byte against the oracle, and the paths the smaller tests leave out:
  * seeded long reads (one 16,500 bp template, all 100 reads, -r0/-r1/-r2);
  * protein families (2,000 families x 16 members, -f0, drawn matrix);
  * the serving stream: jobs of 15 kb templates with 50+ reads and
    incremental jobs whose 50-read graph is loaded from GFA text (-i), all
    arriving while others run, served host-driven and on device-resident
    graphs;
  * the batched workload: all 4,096 jobs, host-driven and device-resident;
  * the arena-overflow recovery (full-width relaunch) in the one-call path
    and in the engine."""
import os
import time

import numpy as np
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, Sequence, records, synth
from test_io_cpu import oracle_copy

pytestmark = pytest.mark.gpu

CORES = min(16, len(os.sched_getaffinity(0)))


def fnv1a(data, h=14695981039346656037):
    for x in data:
        h ^= x
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_seeded_long_read_job_full_depth_matches_oracle(ctx, oracle):
    """16,500 bp template, 100 reads at 10% trimmed by up to 500 bases at
    each end, minimizer windows (-S): consensus + MSA and GFA bytes of the
    whole job equal the oracle's (which runs each read's windows on host
    threads, the reference's -t)."""
    jo, ro, b = synth.seeded(seed=20263, njobs=1)
    assert int(jo[1] - jo[0]) == 100
    p = Params(seeding=True)
    res = mh.batch_run(ctx, p, jo, ro, b)
    assert res.status(0) == 0, mh.last_error()
    cells, want, want_gfa = oracle.batch_run(Params(seeding=True, threads=CORES), jo, ro, b, 1,
                                             gfa=True)
    assert want[0] != 0 and want_gfa[0] != 0
    assert fnv1a((res.output(0, 0) + res.output(0, 1)).encode()) == int(want[0])
    assert fnv1a(res.output(0, 2).encode()) == int(want_gfa[0])
    assert res.stats()["cells"] == cells


def test_protein_families_full_size_match_oracle(ctx, oracle):
    """2,000 protein families (root U[100,800], 16 members at 30%
    divergence), 20-letter alphabet, drawn symmetric matrix through the -m
    file format, full DP: every family's consensus + MSA equals the oracle's,
    on the host-driven engine and device-resident."""
    params, jo, ro, b = synth.protein_families()
    assert len(jo) - 1 == 2000 and all(int(jo[k + 1] - jo[k]) == 16 for k in range(2000))
    import tempfile
    from pathlib import Path

    with tempfile.TemporaryDirectory() as d:
        f = Path(d) / "protein.mtx"
        f.write_text(synth.matrix_text(params.alphabet, params.matrix))
        params = mh.load_matrix(Params(f=0), f)
    cells, want = oracle.batch_run(params, jo, ro, b, CORES)
    # The host-driven engine, then device-resident (the persistent kernel).
    for resident in (False, True):
        res = mh.batch_run(ctx, params, jo, ro, b, resident=resident)
        got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode()) for j in range(2000)],
                       dtype=np.uint64)
        assert res.stats()["cells"] == cells, resident
        assert (got == want).all(), (resident, np.nonzero(got != want))
        if resident:
            assert res.stats()["resident_jobs"] == 2000
        del res


# Jobs of the serving stream (seed 20265): 19 (15,522 bp, 50 reads loaded
# from GFA + 20 added), 84 (15,313 bp, 53 reads), 16 and 53 (incremental,
# 1,129 and 395 bp) and four ordinary jobs.
SERVED = [19, 84, 16, 53, 1, 2, 3, 4]


def _gfa_of(oracle, params, reads):
    """The -r2 text of a graph built from `reads` (reads named read<k>): the
    file a -i run would load."""
    seqs = records(reads)
    og = oracle.align_sequences(params, seqs)
    return og.gfa(params, seqs)


def _render(params, g, seqs):
    return (">consensus sequence\n" + g.consensus(params)[0].decode() + "\n",
            g.msa(params, seqs), g.gfa(params, seqs))


def test_serving_stream_at_size_with_gfa_loaded_incremental_jobs(ctx, oracle):
    """Jobs of the serving workload at their real sizes arrive while others
    run (submitted at their arrival offsets, compressed 10x); the
    incremental ones continue a 50-read graph loaded from GFA text with
    read_gfa, as -i does. Every job's -r0/-r1/-r2 bytes equal the oracle's:
    from scratch for ordinary jobs, and for incremental jobs the oracle
    continues a field-by-field copy of the loaded graph. The oracle runs on
    host threads while the device serves."""
    from concurrent.futures import ThreadPoolExecutor

    p = Params(adaptive=True)
    stream = synth.serving_stream()
    jobs = [stream[k] for k in SERVED]
    assert len(jobs[0]["reads"][0]) > 15000 and len(jobs[0]["base"]) == 50
    assert len(jobs[1]["reads"]) >= 50
    pool = ThreadPoolExecutor(max_workers=CORES)
    texts = {k: pool.submit(_gfa_of, oracle, p, j["base"]) for k, j in enumerate(jobs) if j["base"]}
    loaded = {}
    for k, f in texts.items():
        g, rec = mh.read_gfa(p, f.result())
        assert len(rec) == 50
        loaded[k] = (g, rec, oracle_copy(oracle, p, g))

    def expected(k):
        j = jobs[k]
        if k in loaded:
            _, rec, og = loaded[k]
            seqs = rec + [Sequence(name=f"read{50 + t}", seq=r) for t, r in enumerate(j["reads"])]
            oracle.extend_graph(p, og, seqs, 50)
            return _render(p, og, seqs)
        seqs = records(j["reads"])
        return _render(p, oracle.align_sequences(p, seqs), seqs)

    # The biggest jobs' expectations start first.
    want = {k: pool.submit(expected, k) for k in sorted(range(len(jobs)),
                                                        key=lambda k: -len(jobs[k]["reads"][0]))}
    # Host-driven serving, then serving on device-resident graphs (the
    # GFA-loaded graphs uploaded to the device with their list orders).
    for resident in (False, True):
        print("resident", resident, flush=True)
        with mh.Server(ctx, p, resident=resident) as srv:
            tickets = []
            t0 = time.perf_counter()
            for k, j in enumerate(jobs):
                delay = (j["t"] - jobs[0]["t"]) / 10 - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                if k in loaded:
                    g, rec, _ = loaded[k]
                    tickets.append(srv.submit(j["reads"], graph=g, existing=len(rec)))
                else:
                    tickets.append(srv.submit(j["reads"]))
            got = []
            for t in tickets:
                srv.wait(t)
                assert srv.status(t) == 0, mh.last_error()
                got.append((srv.output(t, 0), srv.output(t, 1), srv.output(t, 2)))
            if resident:
                assert srv.stats()["resident_jobs"] == len(jobs)
        for k in range(len(jobs)):
            assert got[k] == want[k].result(), (resident, SERVED[k])
    pool.shutdown()


def test_batched_workload_full_size_host_and_resident_match_oracle(ctx, oracle):
    """All 4,096 jobs of the batched workload (L in [200, 2000], 8-32 reads
    at 5-15% error, adaptive band), host-driven and device-resident:
    consensus FASTA + MSA hashes of every job equal the oracle's, and so do
    the cells."""
    jo, ro, b = synth.jobs(seed=20262, njobs=4096)
    params = Params(adaptive=True)
    cells, want = oracle.batch_run(params, jo, ro, b, CORES)
    for resident in (False, True):
        print("resident", resident, flush=True)
        res = mh.batch_run(ctx, params, jo, ro, b, resident=resident)
        got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode()) for j in range(4096)],
                       dtype=np.uint64)
        assert (got == want).all(), (resident, np.nonzero(got != want)[0][:10])
        assert res.stats()["cells"] == cells
        del res


def test_gfa_loaded_graph_extends_on_device_as_the_oracle(ctx, oracle):
    """tests/io_cli.rs:563-581 (-i graph + one read, -r1) and a loaded
    graph with branches extended by reads through mpoa_extend_graph."""
    p = Params()
    g, rec = mh.read_gfa(p, "S\ta\tACGT\nP\told\ta+\t*\n")
    seqs = rec + [Sequence(name="new", seq=b"ACGT")]
    mh.extend_graph(ctx, p, g, seqs, len(rec))
    assert g.msa(p, seqs) == ">old \nACGT\n>new \nACGT\n"
    reads = synth.c1(length=600, nreads=9)
    g, rec = mh.read_gfa(p, _gfa_of(oracle, p, reads[:6]))
    og = oracle_copy(oracle, p, g)
    seqs = rec + [Sequence(name=f"read{6 + t}", seq=r) for t, r in enumerate(reads[6:])]
    mh.extend_graph(ctx, p, g, seqs, 6)
    oracle.extend_graph(p, og, seqs, 6)
    assert _render(p, g, seqs) == _render(p, og, seqs)


@pytest.fixture
def squeezed_arena():
    os.environ["MPOA_ARENA_SQUEEZE"] = "1"
    yield
    os.environ.pop("MPOA_ARENA_SQUEEZE", None)


def test_arena_overflow_relaunch_one_call_and_engine(ctx, oracle, squeezed_arena):
    """With 64-element private arenas and a one-chunk pool, DPs overflow
    (the fill's kArenaOverflow) and are re-run with full-width arena rows:
    the one-call path's second pass and the engine's requeue both give the
    oracle's bytes, and the engine counts the relaunches."""
    for params in (Params(adaptive=True), Params()):
        reads = synth.c1(length=500, nreads=6)
        seqs = records(reads)
        dg = mh.align_sequences(ctx, params, seqs)
        og = oracle.align_sequences(params, seqs)
        assert _render(params, dg, seqs) == _render(params, og, seqs)
    jo, ro, b = synth.jobs(seed=77, njobs=48, lmin=200, lmax=900, rmin=4, rmax=10)
    p = Params(adaptive=True)
    res = mh.batch_run(ctx, p, jo, ro, b)
    st = res.stats()
    assert st["arena_retries"] > 0
    cells, want = oracle.batch_run(p, jo, ro, b, CORES)
    got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode()) for j in range(48)],
                   dtype=np.uint64)
    assert (got == want).all(), np.nonzero(got != want)
    assert st["cells"] == cells
