"""The library's host seeding (minipoa_hip.sketch ... anchor_windows, through
the C ABI) against the oracle's restatement of src/minimizer.rs on seeded
random inputs: DNA with ambiguous runs, lower case and U, tandem repeats,
rearranged and diverged copies. CPU only (no device calls)."""
import random

import pytest

import minipoa_hip as mh
from minipoa_hip.params import Params


def mutate(rng, seq, rate):
    out = bytearray()
    for c in seq:
        r = rng.random()
        if r < rate * 0.6:
            out.append(rng.choice(b"ACGT"))
        elif r < rate * 0.8:
            out.append(c)
            out.append(rng.choice(b"ACGT"))
        elif r < rate:
            continue
        else:
            out.append(c)
    return bytes(out)


def random_dna(rng, n, alphabet=b"ACGT"):
    return bytes(rng.choice(alphabet) for _ in range(n))


def with_repeats(rng, seq):
    """Inserts a few tandem repeats (period 2..40, 3..12 copies)."""
    s = bytearray(seq)
    for _ in range(rng.randint(1, 4)):
        unit = random_dna(rng, rng.randint(2, 40))
        at = rng.randrange(len(s) + 1)
        s[at:at] = unit * rng.randint(3, 12)
    return bytes(s)


def messy(rng, seq):
    """Ambiguous runs, lower case and U in place of T."""
    s = bytearray(seq)
    for _ in range(rng.randint(0, 4)):
        at = rng.randrange(len(s) + 1)
        s[at:at] = b"N" * rng.randint(1, 30)
    for i in range(len(s)):
        r = rng.random()
        if r < 0.05:
            s[i] = s[i] | 0x20 if s[i] != ord("N") else s[i]
        elif r < 0.07 and s[i] == ord("T"):
            s[i] = ord("U")
    return bytes(s)


@pytest.mark.parametrize("seed", range(6))
def test_sketch_matches_oracle(oracle, seed):
    rng = random.Random(1000 + seed)
    seq = messy(rng, with_repeats(rng, random_dna(rng, rng.randint(50, 6000))))
    for k, w in [(1, 1), (3, 5), (7, 4), (15, 10), (19, 10), (31, 1), (31, 32767), (11, 200)]:
        assert mh.sketch(seq, k, w) == [mh.Minimizer(*m) for m in oracle.sketch(seq, k, w)], (k, w)


def test_sketch_rejects_like_oracle(oracle):
    for seq, k, w in [(b"ACGT", 0, 2), (b"ACGT", 32, 2), (b"ACGT", 3, 0), (b"ACGT", 3, 32768),
                      (b"", 5, 5), (b"NNNN", 3, 3), (b"AC", 3, 3)]:
        assert mh.sketch(seq, k, w) == [] == oracle.sketch(seq, k, w)


def _params(rng):
    return Params(k=rng.choice([5, 7, 11, 15, 19]), w=rng.choice([1, 3, 4, 10]),
                  tr_dist=rng.choice([0, 30, 1000]), max_occ=rng.choice([0, 1, 3]),
                  win_band=rng.choice([20, 100, 500]), win_max=rng.choice([80, 500, 10000]))


def _pair(rng):
    """(query, target): a diverged copy, sometimes rearranged or trimmed."""
    t = with_repeats(rng, random_dna(rng, rng.randint(300, 8000)))
    q = mutate(rng, t, rng.choice([0.0, 0.02, 0.1, 0.2]))
    r = rng.random()
    if r < 0.25 and len(q) > 100:
        cut = rng.randrange(len(q))
        q = q[cut:] + q[:cut]
    elif r < 0.5:
        q = q[rng.randint(0, len(q) // 4):len(q) - rng.randint(0, len(q) // 4)]
    return q, t


def _anchors(seq):
    return [mh.Anchor(*a) for a in seq]


@pytest.mark.parametrize("seed", range(10))
def test_anchors_chain_and_windows_match_oracle(oracle, seed):
    rng = random.Random(2000 + seed)
    p = _params(rng)
    q, t = _pair(rng)
    got = mh.collect_anchors(p, q, t)
    assert got == _anchors(oracle.collect_anchors(p, q, t))
    ch = mh.chain(got, p.k)
    assert ch == _anchors(oracle.chain([oracle.Anchor(*a) for a in got], p.k))
    mw = min(p.win_max, max(500, len(q) // 25))
    w, tail = mh.partition_windows(ch, len(t), len(q), p.win_band, mw)
    ow, otail = oracle.partition_windows([oracle.Anchor(*a) for a in ch], len(t), len(q),
                                         p.win_band, mw)
    assert (w, tail) == (_anchors(ow), otail)
    aw = mh.anchor_windows(p, q, t)
    ao, at = oracle.anchor_windows(p, q, t)
    assert aw == (_anchors(ao), at)


@pytest.mark.parametrize("seed", range(6))
def test_chain_of_synthetic_anchor_clouds_matches_oracle(oracle, seed):
    """Dense random anchor sets (diagonal runs, noise, equal target ends)
    drive the skip/iteration cut-offs and the local-chain joins."""
    rng = random.Random(3000 + seed)
    pts = set()
    for _ in range(rng.randint(1, 6)):
        d = rng.randint(-300, 300)
        t0 = rng.randint(0, 5000)
        for i in range(rng.randint(3, 400)):
            t = t0 + i * rng.randint(5, 60)
            pts.add((t, max(0, t + d + rng.randint(-3, 3))))
    for _ in range(rng.randint(0, 300)):
        pts.add((rng.randint(0, 20000), rng.randint(0, 20000)))
    anchors = [mh.Anchor(t, q, 19, False) for t, q in sorted(pts) if t >= 18 and q >= 18]
    k = rng.choice([7, 15, 19])
    assert mh.chain(anchors, k) == _anchors(oracle.chain([oracle.Anchor(*a) for a in anchors], k))


def test_partition_windows_edge_bounds_match_oracle(oracle):
    rng = random.Random(4000)
    for _ in range(40):
        n = rng.randint(0, 30)
        t = q = 0
        ch = []
        for _ in range(n):
            t += rng.randint(1, 400)
            q += rng.randint(1, 400)
            ch.append(mh.Anchor(t + 18, q + 18, 19, False))
            t += 19
            q += 19
        tlen, qlen = t + rng.randint(0, 500), q + rng.randint(0, 500)
        drift, win = rng.choice([0, 20, 100, 500]), rng.choice([1, 80, 500, 2000])
        got = mh.partition_windows(ch, tlen, qlen, drift, win)
        ow, ot = oracle.partition_windows([oracle.Anchor(*a) for a in ch], tlen, qlen, drift, win)
        assert got == (_anchors(ow), ot)


def test_protein_alphabet_has_no_anchors():
    """src/minimizer.rs:129-131: more than five symbols, no seeds."""
    import numpy as np

    from minipoa_hip.synth import PROTEIN, protein_matrix
    p = Params(alphabet=PROTEIN, matrix=protein_matrix(np.random.default_rng(1)))
    assert mh.collect_anchors(p, b"ACDEFGHIKL" * 20, b"ACDEFGHIKL" * 20) == []
    # The same text under the DNA alphabet does seed.
    assert mh.collect_anchors(Params(k=5, w=1, tr_dist=0), b"ACGTTGCAAC" * 3, b"ACGTTGCAAC" * 3)


@pytest.mark.parametrize("seed", range(5))
def test_guide_order_matches_oracle(oracle, seed):
    """src/minimizer.rs:489-578: reads of three diverged templates, shuffled,
    with ambiguous runs and one short read without minimizers."""
    from minipoa_hip.params import Sequence
    rng = random.Random(5000 + seed)
    templates = [random_dna(rng, rng.randint(400, 3000)) for _ in range(3)]
    seqs = [Sequence(seq=messy(rng, mutate(rng, rng.choice(templates), rng.choice([0.02, 0.1]))))
            for _ in range(rng.randint(2, 14))]
    seqs.append(Sequence(seq=b"ACG"))
    rng.shuffle(seqs)
    for p in (Params(progressive=True), Params(progressive=True, k=11, w=4),
              Params(progressive=True, k=15, w=5, mm_filter_ratio=0.5),
              Params(progressive=False), Params(progressive=True, mm_filter_ratio=0.0)):
        assert mh.guide_order(p, seqs) == oracle.guide_order(p, seqs), p


def test_guide_order_reference_cases():
    """src/minimizer.rs:753-776 on the library."""
    from minipoa_hip.params import Sequence
    from test_seed_oracle import xorshift_prefix_records
    seqs = xorshift_prefix_records()
    p = Params(progressive=True)
    order = mh.guide_order(p, seqs)
    assert order == mh.guide_order(p, seqs)
    assert sorted(order) == list(range(12))
    assert mh.guide_order(p, [Sequence(seq=b"ACGT")] * 3) == [0, 1, 2]
