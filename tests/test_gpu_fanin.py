"""Rows with many predecessors on the device, against the oracle: a node X
joined by k parallel one-base branches, k = 3 (the predecessor list in one
gather), 70 (two 64-lane gather groups and the one-at-a-time adaptive fold)
and 600 (records of 8 bytes: the wide-record kernel). Repeated edges give
the branches weights 1-3, so `choose` meets both heavier and tied
candidates; a second fan-in node Y, whose branches have bands of different
lengths, makes predecessors whose band ends a column before the current
one. Static, adaptive and full bands, i32 and i64."""
import random

import pytest

import minipoa_hip as mh
from minipoa_hip import Params
from test_io_cpu import oracle_copy

pytestmark = pytest.mark.gpu

ACGT = b"ACGT"


def fan_graph(p, k, seed):
    """source -> 20-base spine -> k branches (one base each) -> X -> 12-base
    spine -> k/2 branches of 1-3 bases -> Y -> 12-base spine -> sink.
    Returns the graph and the bases along one route (for queries)."""
    rng = random.Random(seed)
    g = mh.Graph(p)
    route = []

    def spine(prev, n):
        for _ in range(n):
            b = rng.randrange(4)
            v = g.add_node(b)
            g.add_edge(prev, v)
            route.append(b)
            prev = v
        return prev

    def fan(prev, n, maxlen):
        ends = []
        for _ in range(n):
            u = prev
            for _ in range(rng.randint(1, maxlen)):
                v = g.add_node(rng.randrange(4))
                g.add_edge(u, v)
                u = v
            ends.append(u)
        x = g.add_node(rng.randrange(4))
        for u in ends:
            for _ in range(rng.choice((1, 1, 2, 3))):
                g.add_edge(u, x)
        route.append(rng.randrange(4))  # a base standing for the branch taken
        route.append(g.view().base[x])
        return x

    last = spine(0, 20)
    last = fan(last, k, 1)
    last = spine(last, 12)
    last = fan(last, max(2, k // 2), 3)
    last = spine(last, 12)
    g.add_edge(last, 1)
    g.toposort(False)
    return g, route


def queries(route, seed):
    rng = random.Random(seed)
    out = [bytes(ACGT[b] for b in route)]
    for _ in range(3):
        q = bytearray()
        for b in route:
            x = rng.random()
            if x < 0.06:
                continue
            q.append(rng.choice(ACGT) if x < 0.14 else ACGT[b])
            if rng.random() < 0.05:
                q.append(rng.choice(ACGT))
        out.append(bytes(q))
    return out


@pytest.mark.parametrize("k", [3, 70, 600])
def test_fan_in_rows_match_oracle(ctx, oracle, k):
    """Paths, errors and final scores of both device kernels equal the
    oracle's i64 kernel (src/align.rs:531-545)."""
    for pi, p in enumerate((Params(), Params(adaptive=True), Params(f=0))):
        g, route = fan_graph(p, k, seed=1000 * k + pi)
        og = oracle_copy(oracle, p, g)
        for qi, q in enumerate(queries(route, seed=k + pi)):
            for adaptive, band in ((p.adaptive, p.b), (True, 4), (False, 3)):
                ref = oracle.result_or_error(oracle.poa_scalar_i64, p, og, 0, 1, q, adaptive, band)
                ref_score = oracle.last_score() if ref[0] == "ok" else None
                for bits in (32, 64):
                    got = mh.result_or_error(mh.poa_kernel, ctx, p, g, 0, 1, q, adaptive, band, bits,
                                             score=True)
                    case = (k, pi, qi, adaptive, band, bits)
                    if ref[0] == "ok":
                        assert got[0] == "ok", (case, got)
                        assert got[1][0] == ref[1], case
                        assert got[1][1] == ref_score, case
                    else:
                        assert got == ref, case
