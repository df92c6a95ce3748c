"""Batched progressive POA on the device: every job's consensus / MSA / GFA
bytes equal the oracle's for the same reads, alone or in a batch."""
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


def test_small_batch_matches_oracle_bytes(ctx, oracle):
    jo, ro, b = synth.jobs(seed=7, njobs=24, lmin=40, lmax=300, rmin=2, rmax=7)
    for params in (Params(), Params(adaptive=True), Params(f=0), Params(b=2, adaptive=True)):
        res = mh.batch_run(ctx, params, jo, ro, b)
        for j in range(len(jo) - 1):
            seqs = records(synth.job_reads(jo, ro, b, j))
            og = oracle.align_sequences(params, seqs)
            assert res.status(j) == 0, mh.last_error()
            assert res.output(j, 0) == ">consensus sequence\n" + og.consensus(params)[0].decode() + "\n"
            assert res.output(j, 1) == og.msa(params, seqs)
            assert res.output(j, 2) == og.gfa(params, seqs)


def test_batch_equals_solo_device_runs(ctx):
    """A job's result is the same alone (one DP per launch) and in the mix."""
    jo, ro, b = synth.jobs(seed=11, njobs=12, lmin=100, lmax=400, rmin=3, rmax=6)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b)
    for j in range(len(jo) - 1):
        seqs = records(synth.job_reads(jo, ro, b, j))
        g = mh.align_sequences(ctx, params, seqs)
        assert res.output(j, 1) == g.msa(params, seqs)
        assert res.output(j, 2) == g.gfa(params, seqs)


def test_batched_workload_sample_matches_oracle_hashes(ctx, oracle):
    """A 64-job sample of the 4,096-job batched workload (L in [200, 2000],
    8-32 reads at 5-15% error, adaptive band): FNV-1a of consensus FASTA +
    MSA per job equals the oracle's."""
    jo, ro, b = synth.jobs(seed=20262, njobs=64)
    params = Params(adaptive=True)
    res = mh.batch_run(ctx, params, jo, ro, b)
    cells, want = oracle.batch_run(params, jo, ro, b, 8)
    got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode()) for j in range(64)],
                   dtype=np.uint64)
    assert (want != 0).all()
    assert (got == want).all(), np.nonzero(got != want)
    assert res.stats()["cells"] == cells


def test_records_layout_needs_less_device_memory_same_bytes(oracle):
    """The same batch on a values context and on a records context: equal
    outputs, and the records context's device buffers (arena included) are
    smaller."""
    jo, ro, b = synth.jobs(seed=31, njobs=24, lmin=300, lmax=1500, rmin=4, rmax=12)
    p = Params(adaptive=True)
    out, size = {}, {}
    for name, lay in (("values", mh.LAYOUT_VALUES), ("records", mh.LAYOUT_RECORDS)):
        c = mh.Context(0)
        c.set_layout(lay)
        res = mh.batch_run(c, p, jo, ro, b)
        out[name] = [(res.output(j, 0), res.output(j, 1)) for j in range(24)]
        size[name] = c.device_bytes()
        c.close()
    assert out["values"] == out["records"]
    assert size["records"] < size["values"], size
