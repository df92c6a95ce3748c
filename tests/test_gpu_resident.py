"""Device-resident batches (mpoa_batch_run_resident): every job's graph and
DP descriptors stay on the device between reads, fusion and planning run
on the device, either in one launch of the persistent kernel (the default:
each workgroup runs jobs from a device queue to their end) or as length
buckets whose step loop is a replayed HIP graph (MPOA_RES_PERSIST=0). Every
test runs in both modes. Outputs must be byte-identical to the oracle's and
to the host-driven batch, including for jobs that leave the device
part-way."""
import os

import numpy as np
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, records, synth

pytestmark = pytest.mark.gpu


def fnv1a(data, h=14695981039346656037):
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def pack(job_reads):
    lens = [len(r) for reads in job_reads for r in reads]
    jo = np.zeros(len(job_reads) + 1, dtype=np.uint64)
    jo[1:] = np.cumsum([len(reads) for reads in job_reads])
    ro = np.zeros(len(lens) + 1, dtype=np.uint64)
    ro[1:] = np.cumsum(lens)
    b = np.frombuffer(b"".join(r for reads in job_reads for r in reads) or b"\0", dtype=np.uint8)
    return jo, ro, b


def check_against_oracle(res, oracle, params, jo, ro, b, gfa=True):
    for j in range(len(jo) - 1):
        seqs = records(synth.job_reads(jo, ro, b, j))
        og = oracle.align_sequences(params, seqs)
        assert res.status(j) == 0, (j, mh.last_error())
        assert res.output(j, 0) == ">consensus sequence\n" + og.consensus(params)[0].decode() + "\n", j
        assert res.output(j, 1) == og.msa(params, seqs), j
        if gfa:
            assert res.output(j, 2) == og.gfa(params, seqs), j


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


@pytest.fixture(params=["persistent", "captured"], autouse=True)
def mode(request, env):
    env("MPOA_RES_PERSIST", "1" if request.param == "persistent" else "0")
    print("resident mode", request.param, flush=True)
    return request.param


def test_resident_small_batches_match_oracle(ctx, oracle):
    jo, ro, b = synth.jobs(seed=7, njobs=24, lmin=40, lmax=300, rmin=2, rmax=7)
    for params in (Params(), Params(adaptive=True), Params(f=0), Params(b=2, adaptive=True)):
        print("params", params.adaptive, params.f, params.b, flush=True)
        res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
        st = res.stats()
        assert st["resident_jobs"] + st["evicted_jobs"] == 24 and st["step_replays"] > 0
        check_against_oracle(res, oracle, params, jo, ro, b)


def test_resident_batched_workload_sample_matches_oracle_and_host(ctx, oracle):
    """64 jobs of the 4,096-job workload (L in [200, 2000], 8-32 reads at
    5-15% error, adaptive band): consensus + MSA hashes equal the oracle's,
    GFA bytes equal the host-driven batch's, and so do the cells."""
    jo, ro, b = synth.jobs(seed=20262, njobs=64)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    st = res.stats()
    cells, want = oracle.batch_run(params, jo, ro, b, 8)
    got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode()) for j in range(64)],
                   dtype=np.uint64)
    assert (got == want).all(), np.nonzero(got != want)
    assert st["cells"] == cells
    assert st["resident_jobs"] == 64 and st["evicted_jobs"] == 0
    host = mh.batch_run(ctx, params, jo, ro, b)
    for j in range(64):
        assert res.output(j, 2) == host.output(j, 2), j
    assert st["pred_cells"] == host.stats()["pred_cells"]


def test_resident_many_buckets_over_all_slots(ctx, oracle, env, mode):
    """Buckets of three jobs: many step graphs, every launch slot busy."""
    if mode == "persistent":
        pytest.skip("one launch holds every job")
    env("MPOA_RES_BUCKET", "3")
    jo, ro, b = synth.jobs(seed=13, njobs=20, lmin=60, lmax=700, rmin=1, rmax=9)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    assert res.stats()["buckets"] == 7
    check_against_oracle(res, oracle, params, jo, ro, b)


def test_resident_edge_cases(ctx, oracle):
    """Empty reads (first and later), a single read, single bases, all-N
    reads, lower case and RNA letters, and a job with no reads (host path)."""
    jobs = [
        [b"", b"ACGT", b"", b"ACGA"],
        [b"ACGTACGTTT"],
        [b"A", b"C", b"A", b"G"],
        [b"NNNNNN", b"ACGTNN", b"NNNN"],
        [b"acgtacgtaa", b"ACGUACGUAA", b"acguacgtaa"],
        [],
        [b"", b""],
        [b"GATTACA" * 30, b"GATTACA" * 29 + b"GATTAC", b"GATTACAGATTACA"],
    ]
    jo, ro, b = pack(jobs)
    for params in (Params(adaptive=True), Params(f=0)):
        res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
        check_against_oracle(res, oracle, params, jo, ro, b)
        st = res.stats()
        assert st["resident_jobs"] == len(jobs) - 1  # the job without reads runs on the host


def test_resident_jobs_leave_the_device_when_the_score_guard_fails(ctx, oracle):
    """Scores whose products pass 2^26 as the graph grows
    (src/align.rs:493-529): the device plans until the guard fails, then the
    job continues on the host path (i64 kernel) from that read."""
    # |score| 2^16: a DP is i32-safe while rows + query + 20 <= 1,024.
    p = Params(match_score=1 << 16, mismatch=-(1 << 16), gap_open=-(1 << 16), gap_ext=-(1 << 15)).rebuild_matrix()
    jo, ro, b = synth.jobs(seed=17, njobs=10, lmin=300, lmax=700, rmin=3, rmax=8)
    res = mh.batch_run(ctx, p, jo, ro, b, resident=True)
    st = res.stats()
    assert st["evicted_jobs"] > 0 and st["i64_calls"] > 0
    check_against_oracle(res, oracle, p, jo, ro, b)


def test_resident_jobs_leave_the_device_when_the_arena_is_too_small(ctx, oracle, env):
    """A 1 MiB device arena: DPs that do not fit leave the device and finish
    on the host engine, with the same bytes."""
    env("MPOA_RES_ARENA_MB", "1")
    jo, ro, b = synth.jobs(seed=19, njobs=12, lmin=100, lmax=900, rmin=2, rmax=8)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    st = res.stats()
    assert st["evicted_jobs"] > 0
    check_against_oracle(res, oracle, params, jo, ro, b)


def test_resident_persistent_dps_extend_into_the_shared_pool(ctx, oracle, env, mode):
    """Workgroup regions of 1 MiB and a 512 MiB shared pool: DPs outgrow
    their region and take pool chunks instead of leaving the device."""
    if mode == "captured":
        pytest.skip("regions belong to the persistent kernel")
    env("MPOA_RES_ARENA_MB", "1")
    env("MPOA_RES_POOL_MB", "512")
    jo, ro, b = synth.jobs(seed=19, njobs=12, lmin=100, lmax=900, rmin=2, rmax=8)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    st = res.stats()
    assert st["evicted_jobs"] == 0 and st["resident_jobs"] == 12
    check_against_oracle(res, oracle, params, jo, ro, b)


def test_resident_modes_agree(ctx, env, mode):
    """Persistent and captured runs of 96 workload jobs: same bytes, cells
    and band overlaps."""
    if mode == "captured":
        pytest.skip("compared within the persistent case")
    jo, ro, b = synth.jobs(seed=20263, njobs=96)
    params = Params(adaptive=True)
    a = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    env("MPOA_RES_PERSIST", "0")
    c = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    for j in range(96):
        assert (a.output(j, 0), a.output(j, 1), a.output(j, 2)) == (c.output(j, 0), c.output(j, 1), c.output(j, 2)), j
    sa, sc = a.stats(), c.stats()
    assert sa["cells"] == sc["cells"] and sa["pred_cells"] == sc["pred_cells"] and sa["dp_calls"] == sc["dp_calls"]
    assert sa["buckets"] == 1 and sc["buckets"] > 1


@pytest.mark.parametrize("scap", ["0", "2000000", "-"])
def test_resident_streamed_and_after_launch_jobs_agree(ctx, oracle, env, mode, scap):
    """Persistent launch: finished jobs stream back while it runs; with the
    streaming share capped (none, about 2 MB: some jobs stream, the rest are
    packed after the launch) or streaming off, every job's bytes equal the
    oracle's and the uncapped run's, and the statistics agree."""
    if mode == "captured":
        pytest.skip("streaming belongs to the persistent launch")
    jo, ro, b = synth.jobs(seed=20265, njobs=48, lmin=200, lmax=1200, rmin=3, rmax=12)
    params = Params(adaptive=True)
    a = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    if scap == "-":
        env("MPOA_PERSIST_STREAM", "0")
    else:
        env("MPOA_PERSIST_SCAP", scap)
    c = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    for j in range(48):
        assert (a.output(j, 0), a.output(j, 1), a.output(j, 2)) == (c.output(j, 0), c.output(j, 1), c.output(j, 2)), j
    sa, sc = a.stats(), c.stats()
    for k in ("cells", "pred_cells", "dp_calls", "resident_jobs", "evicted_jobs"):
        assert sa[k] == sc[k], k
    check_against_oracle(c, oracle, params, jo, ro, b)


def test_resident_wide_rows_and_long_reads_match_oracle(ctx, oracle):
    """Templates of 3,000-12,500 bases: rows wider than one 4x4 stripe (1,024
    columns) and query codes of up to 12.5 KB in LDS, which shrink the fill's
    ring and (persistent kernel) the workgroups per CU."""
    rng = np.random.Generator(np.random.PCG64(31))
    jobs = []
    for L, R in ((3000, 4), (6000, 3), (12500, 2)):
        tmpl = rng.integers(0, 4, size=L).astype(np.uint8)
        jobs.append([synth.mutate(rng, tmpl, 0.08) for _ in range(R)])
    jo, ro, b = pack(jobs)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    st = res.stats()
    assert st["resident_jobs"] == 3 and st["evicted_jobs"] == 0
    check_against_oracle(res, oracle, params, jo, ro, b)


def test_resident_mixed_small_and_long_jobs_match_oracle(ctx, oracle):
    """Forty small jobs beside two long ones in one batch: the persistent
    kernel's arena regions are sized by the longest template, and the queue
    takes the long jobs first."""
    small = synth.jobs(seed=37, njobs=40, lmin=60, lmax=300, rmin=2, rmax=6)
    big = synth.jobs(seed=38, njobs=2, lmin=3500, lmax=4000, rmin=3, rmax=4)
    reads = [synth.job_reads(*small, j) for j in range(40)] + [synth.job_reads(*big, j) for j in range(2)]
    jo, ro, b = pack(reads)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
    st = res.stats()
    assert st["resident_jobs"] == 42 and st["evicted_jobs"] == 0
    check_against_oracle(res, oracle, params, jo, ro, b, gfa=False)


def test_resident_seeded_and_guide_order_jobs_run_on_the_host(ctx, oracle):
    """Seeding and the -p guide order stay on the host-driven engine."""
    jo, ro, b = synth.jobs(seed=23, njobs=4, lmin=300, lmax=600, rmin=3, rmax=5)
    for params in (Params(seeding=True, k=11, w=4, win_band=30, win_max=100), Params(progressive=True)):
        res = mh.batch_run(ctx, params, jo, ro, b, resident=True)
        assert res.stats()["resident_jobs"] == 0
        host = mh.batch_run(ctx, params, jo, ro, b)
        for j in range(4):
            assert res.output(j, 1) == host.output(j, 1)
            assert res.output(j, 2) == host.output(j, 2)
