"""mpoa_poa_noalloc: the same DP as mpoa_poa with every device buffer taken
from the library's fixed per-GPU scratch. Its paths, errors and scores
must equal mpoa_poa's (and the oracle's) on consecutive calls that reuse the
scratch, the second call smaller than the first."""
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, records, synth

pytestmark = pytest.mark.gpu


def _graph(ctx, params, reads):
    return mh.align_sequences(ctx, params, records(reads))


def _both(ctx, params, g, query, adaptive, band):
    want = mh.result_or_error(mh.poa, ctx, params, g, 0, 1, query, adaptive, band, score=True)
    got = mh.result_or_error(mh.poa_noalloc, ctx, params, g, 0, 1, query, adaptive, band,
                             score=True)
    return want, got


def test_noalloc_matches_poa_on_consecutive_calls_large_then_small(ctx, oracle):
    """A 2 kb, 9-read graph against a tenth read (the plain consensus
    workload's last DP), then a 300 bp graph: each call equals mpoa_poa and
    the oracle, static and adaptive bands."""
    big = synth.c1()  # 2,000 bp template, 10 reads at 10% error
    small = synth.c1(length=300, nreads=4)
    for adaptive in (False, True):
        p = Params(adaptive=adaptive)
        for reads in (big, small):
            g = _graph(ctx, p, reads[:-1])
            query = reads[-1]
            want, got = _both(ctx, p, g, query, adaptive, p.b)
            assert want[0] == "ok", want
            assert got == want, (adaptive, len(query))
            og = oracle.align_sequences(p, records(reads[:-1]))
            ref = oracle.result_or_error(oracle.poa, p, og, 0, 1, query, adaptive, p.b)
            assert got[1][0] == ref[1]


def test_noalloc_band_failure_retry_and_i64_route(ctx, oracle):
    """poa's band -> full-DP retry (band 0 leaves the band) and the i64
    route of extreme scores both run inside the scratch, as mpoa_poa."""
    reads = synth.c1(length=400, nreads=5)
    p = Params(b=0, f=40)
    g = _graph(ctx, Params(), reads[:-1])
    for band in (0, 1):
        want, got = _both(ctx, p, g, reads[-1], False, band)
        assert got == want
    extreme = Params().clone(matrix=[2**31 - 1] * 25)
    assert not mh.i32_scores_safe(extreme, g, 0, 1, len(reads[-1]))
    want, got = _both(ctx, extreme, g, reads[-1][:60], False, 100)
    assert got == want


def test_noalloc_rejects_a_dp_beyond_the_scratch_then_runs_again(ctx):
    """A full-DP matrix of ~6,000 x 6,000 cells exceeds the 64 MiB scratch
    (in the records layout its rows, too wide for the LDS ring, keep M and D:
    ~330 MB): a clear status, and the next smaller call is still equal to
    mpoa_poa."""
    reads = synth.c1(length=6000, nreads=2)
    p = Params(f=0)
    g = _graph(ctx, p, reads[:1])
    with pytest.raises(mh.HipPoaError) as e:
        mh.poa_noalloc(ctx, p, g, 0, 1, reads[1], False, 100)
    assert e.value.status == mh.SCRATCH_LIMIT
    assert "no-allocation scratch" in str(e.value)
    small = synth.c1(length=250, nreads=3)
    g2 = _graph(ctx, Params(), small[:-1])
    want, got = _both(ctx, Params(), g2, small[-1], False, 100)
    assert got == want and got[0] == "ok"
