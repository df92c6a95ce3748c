"""Random small DAGs (2-7 nodes with 1-3 predecessors each, edges of weight
1-3) against random queries of 1-12 bases under tiny static or adaptive
bands (0-4) and random scores: the device kernels (i32 and i64) equal the
oracle's i64 kernel in path, error and score. Tiny bands leave the rows'
bands of different lengths, so the traceback often meets a predecessor
whose band ends one column before the cell (about 1 case in 120 here has a
path that depends on it)."""
import random

import pytest

import minipoa_hip as mh
from minipoa_hip import Params
from rule_kats import Case

pytestmark = pytest.mark.gpu


def small_case(seed):
    rng = random.Random(seed)
    nb = rng.randint(2, 7)
    bases = [rng.randrange(4) for _ in range(nb)]
    ids = list(range(2, 2 + nb))
    edges = []
    for k, v in enumerate(ids):
        cands = [0] + ids[:k]
        for u in rng.sample(cands, min(len(cands), rng.randint(1, 3))):
            edges += [(u, v)] * rng.choice((1, 1, 2, 3))
    has_out = {u for u, _ in edges}
    for v in ids:
        if v not in has_out or rng.random() < 0.3:
            edges += [(v, 1)] * rng.choice((1, 2))
    q = bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 12)))
    p = Params(gap_open=rng.choice((0, -1, -2, -4)), gap_ext=rng.choice((-1, -2)),
               match_score=rng.choice((1, 2)), mismatch=rng.choice((-1, -2, -4)), f=1000)
    p = p.rebuild_matrix()
    band = rng.randint(0, 4)
    adaptive = rng.random() < 0.3
    return Case(f"small{seed}", p, bases, edges, q, adaptive=adaptive, band=band)


def test_small_dags_tiny_bands_match_oracle(ctx, oracle):
    checked = 0
    for seed in range(2000):
        case = small_case(seed)
        og = case.build(oracle.Graph)
        g = case.build(mh.Graph)
        ref = oracle.result_or_error(oracle.poa_scalar_i64, case.params, og, 0, 1, case.query,
                                     case.adaptive, case.band)
        ref_score = oracle.last_score() if ref[0] == "ok" else None
        for bits in (32, 64):
            got = mh.result_or_error(mh.poa_kernel, ctx, case.params, g, 0, 1, case.query,
                                     case.adaptive, case.band, bits, score=True)
            if ref[0] == "ok":
                assert got[0] == "ok", (seed, bits, got)
                assert got[1][0] == ref[1], (seed, bits)
                assert got[1][1] == ref_score, (seed, bits)
            else:
                assert got == ref, (seed, bits)
            checked += 1
    assert checked == 4000
