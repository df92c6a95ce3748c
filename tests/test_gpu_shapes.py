"""Every compiled fill launch shape (waves per job x columns per lane) gives
the same paths, errors, scores and output bytes as the oracle.

The shape is forced through MPOA_FILL_SHAPE, which the runtime reads at
every launch. Cases: src/align.rs:1603-1669 (branched graphs, i32 and i64),
src/align.rs:1802-1871 (1,080 lane/band edge cases), and a sample of the
4,096-job batched workload whose rows span one to several stripes.
"""
import os

import numpy as np
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, records, synth
from test_gpu_batch import fnv1a
from test_gpu_parity import gv, linear
from test_oracle import BAND_SETTINGS, SETTINGS_1080, _branched_cases, cases_1080

pytestmark = pytest.mark.gpu

SHAPES = [(8, 1), (16, 1), (4, 2), (8, 2), (4, 4), (8, 4)]


@pytest.fixture
def shape(request):
    cols, waves = request.param
    old = os.environ.get("MPOA_FILL_SHAPE")
    os.environ["MPOA_FILL_SHAPE"] = f"{cols},{waves}"
    print(f"fill shape {cols},{waves}", flush=True)
    yield request.param
    if old is None:
        del os.environ["MPOA_FILL_SHAPE"]
    else:
        os.environ["MPOA_FILL_SHAPE"] = old


@pytest.mark.parametrize("shape", SHAPES, indirect=True, ids=lambda s: f"c{s[0]}w{s[1]}")
def test_shape_branched_graphs_i32_i64(ctx, oracle, shape):
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
                for bits in (32, 64):
                    got = mh.result_or_error(mh.poa_kernel, ctx, params, view, 0, 1, query,
                                             adaptive, band, bits, score=True)
                    if ref[0] == "ok":
                        assert got[0] == "ok", (shape, case, rid, f, adaptive, band, bits, got)
                        assert got[1][0] == ref[1], (shape, case, rid, f, adaptive, band, bits)
                        assert got[1][1] == ref_score, (shape, case, rid, f, adaptive, band, bits)
                    else:
                        assert got == ref, (shape, case, rid, f, adaptive, band, bits)
                    count += 1
            params.f = 0
            og.add_path(rid, oracle.poa_scalar_i64(params, og, 0, 1, query, False, 100))
            og.toposort(False)
    assert count == 35 * 4 * 4 * 2


@pytest.mark.parametrize("shape", SHAPES, indirect=True, ids=lambda s: f"c{s[0]}w{s[1]}")
def test_shape_lane_and_band_edges(ctx, oracle, shape):
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
                assert got == ref, (shape, length, len(query), f, adaptive, band)
                count += 1
    assert count == 1080


@pytest.fixture(scope="module")
def workload_sample(oracle):
    jo, ro, b = synth.jobs(seed=20262, njobs=24)
    out = {}
    for name, params in (("adaptive", Params(adaptive=True)), ("full", Params(f=0))):
        cells, want = oracle.batch_run(params, jo, ro, b, 8)
        out[name] = (params, cells, want)
    return jo, ro, b, out


@pytest.mark.parametrize("shape", SHAPES, indirect=True, ids=lambda s: f"c{s[0]}w{s[1]}")
def test_shape_batched_workload_sample(ctx, workload_sample, shape):
    """24 jobs of the batched DNA workload (L in [200, 2000]) with the
    adaptive band and with full rows (-f0): consensus + MSA hashes and the
    cell count equal the oracle's."""
    jo, ro, b, out = workload_sample
    for name, (params, cells, want) in out.items():
        res = mh.batch_run(ctx, params, jo, ro, b)
        got = np.array([fnv1a((res.output(j, 0) + res.output(j, 1)).encode())
                        for j in range(len(jo) - 1)], dtype=np.uint64)
        assert (got == want).all(), (shape, name, np.nonzero(got != want))
        assert res.stats()["cells"] == cells, (shape, name)


def test_global_table_kernel_large_alphabet(ctx, oracle):
    """A 30-letter alphabet: its 900-entry score matrix does not fit the LDS
    table, so the fill reads the matrix and query codes from global memory
    (the one-shape fallback kernel). Consensus, MSA and GFA bytes equal the
    oracle's for full, static and adaptive bands."""
    import random

    rng = random.Random(3030)
    alphabet = bytes(range(ord("A"), ord("A") + 26)) + b"0123"
    n = len(alphabet)
    matrix = [0] * (n * n)
    for a in range(n):
        for b in range(a, n):
            v = rng.randint(4, 11) if a == b else rng.randint(-4, 3)
            matrix[a * n + b] = matrix[b * n + a] = v
    root = bytes(rng.choice(alphabet) for _ in range(160))
    reads = []
    for _ in range(6):
        r = bytearray()
        for c in root:
            x = rng.random()
            if x < 0.05:
                continue
            r.append(rng.choice(alphabet) if x < 0.15 else c)
            if rng.random() < 0.05:
                r.append(rng.choice(alphabet))
        reads.append(bytes(r))
    seqs = records(reads)
    for f, adaptive in ((0, False), (40, False), (40, True)):
        params = Params(alphabet=alphabet, matrix=matrix, f=f, adaptive=adaptive)
        dg = mh.align_sequences(ctx, params, seqs)
        og = oracle.align_sequences(params, seqs)
        assert dg.consensus(params) == og.consensus(params), (f, adaptive)
        assert dg.msa(params, seqs) == og.msa(params, seqs), (f, adaptive)
        assert dg.gfa(params, seqs) == og.gfa(params, seqs), (f, adaptive)
