"""Pins the oracle's minimizer seeding and seeded alignment (src/minimizer.rs,
src/align.rs:99-188). CPU only.

Reference tests restated here: src/minimizer.rs:595-784 (all but the
comparison with the simd-minimizers crate's own scalar mode, which is not
vendored), src/align.rs:1525-1559 and tests/avx512.rs:145-180 (seeded
workflows). The crate's ranking of k-mers is not available, so minimizer
POSITIONS are parity unpinned against the reference; here the oracle's
rolling hash and sliding minimum are checked against a direct definition
of the same ranking (every window's k-mers hashed from scratch).
"""
import pytest

from helpers import Lcg, dna, validate_gfa, validate_msa
from minipoa_hip.params import Params, records

NT_SEED = [0x95C60474, 0x62A02B4C, 0x4BE24456, 0x82572324]  # codes A, C, T, G
M32 = 0xFFFFFFFF


def xorshift_dna(n):
    """src/minimizer.rs:593-604"""
    x, out = 42, bytearray()
    for _ in range(n):
        x ^= (x << 13) & (2 ** 64 - 1)
        x ^= x >> 7
        x ^= (x << 17) & (2 ** 64 - 1)
        out.append(b"ACGT"[x & 3])
    return bytes(out)


def rotl(x, r):
    r %= 32
    return ((x << r) | (x >> (32 - r))) & M32 if r else x


def direct_minimizers(seq, k, w):
    """Every window of w k-mers ranked from scratch: key = top 16 bits of the
    forward ntHash XOR_i rotl(seed[c_i], k-1-i); leftmost minimum; distinct
    consecutive picks. One DNA run, len >= k."""
    codes = [(c >> 1) & 3 for c in seq]
    nk = len(seq) - k + 1
    keys = []
    for t in range(nk):
        h = 0
        for i in range(k):
            h ^= rotl(NT_SEED[codes[t + i]], k - 1 - i)
        keys.append(h >> 16)
    w = min(w, nk)
    out, last = [], None
    for s in range(nk - w + 1):
        win = keys[s:s + w]
        pos = s + win.index(min(win))
        if pos != last:
            out.append(pos)
            last = pos
    return out


@pytest.mark.parametrize("n", [1, 5, 19, 31, 128, 1024])
def test_rolling_sketch_matches_direct_ranking(oracle, n):
    """Shape of src/minimizer.rs:606-624, against the direct definition."""
    seq = xorshift_dna(n)
    for k in [k for k in (1, 3, 5, 19, 28, 31) if k <= n]:
        for w in (1, 3, 10, 255, 32767):
            w = min(w, n - k + 1)
            got = [m.end + 1 - k for m in oracle.sketch(seq, k, w)]
            assert got == direct_minimizers(seq, k, w), (n, k, w)


def test_large_window_matches_direct_ranking(oracle):
    """Shape of src/minimizer.rs:640-651 (the largest window), smaller input."""
    seq = xorshift_dna(3000)
    got = [m.end + 1 - 31 for m in oracle.sketch(seq, 31, 32767)]
    assert got == direct_minimizers(seq, 31, 32767)
    assert len(got) >= 1


def test_ambiguous_runs_do_not_create_spanning_seeds(oracle):
    """src/minimizer.rs:626-639"""
    seeds = oracle.sketch(b"acgtACGTNNNNacguACGU", 4, 3)
    assert seeds
    assert all(s.end < 8 or s.end >= 15 for s in seeds)
    assert oracle.sketch(b"acgu", 4, 10) == oracle.sketch(b"ACGT", 4, 10)
    assert oracle.sketch(b"NNNN", 3, 3) == []
    assert oracle.sketch(b"AC", 3, 3) == []
    assert oracle.sketch(b"ACGT", 0, 2) == []
    assert oracle.sketch(b"ACGT", 3, 32768) == []
    assert oracle.sketch(b"ACGT", 3, 2 ** 64 - 1) == []


def test_exact_kmer_values_match(oracle):
    """src/minimizer.rs:653-677"""
    seq = xorshift_dna(1024)
    seq = seq + seq
    bits = {ord("A"): 0, ord("C"): 1, ord("T"): 2, ord("G"): 3}
    for k in (1, 3, 19, 31):
        seeds = oracle.sketch(seq, k, 1)
        assert len(seeds) == len(seq) - k + 1
        seen = {}
        for s in seeds:
            kmer = seq[s.end + 1 - s.span:s.end + 1]
            exact = 0
            for i, b in enumerate(kmer):
                exact |= bits[b] << (2 * i)
            assert s.value == exact, (k, s.end)
            assert seen.setdefault(s.value, kmer) == kmer


def test_repeat_filter_removes_homopolymer_matches(oracle):
    """src/minimizer.rs:679-694"""
    p = Params(k=5, w=3, tr_dist=100)
    a = b"A" * 100
    assert oracle.collect_anchors(p, a, a) == []
    assert oracle.collect_anchors(Params(k=5, w=3, tr_dist=0, max_occ=1), a, a) == []
    # Without either filter the homopolymer anchors every k-mer pair.
    assert oracle.collect_anchors(Params(k=5, w=3, tr_dist=0, max_occ=0), a, a)


def test_repeat_intervals_extend_clip_and_merge(oracle):
    """src/minimizer.rs:696-724"""
    mm = [oracle.Minimizer(1, 9, 5), oracle.Minimizer(1, 19, 5),
          oracle.Minimizer(2, 80, 5), oracle.Minimizer(2, 90, 5)]
    assert oracle.repeat_intervals(mm, 100, 5) == []
    assert oracle.repeat_intervals(mm, 100, 100) == [(0, 100)]
    # Without the 50-base merge gap the two repeats stay apart: periods 10
    # extend [5-... , 19+1+10) and [76-10, 91+10) clipped to 100.
    assert oracle.repeat_intervals(mm[:2], 100, 100) == [(0, 30)]
    assert oracle.repeat_intervals(mm[2:], 100, 100) == [(66, 100)]


def test_chain_and_windows_are_monotone(oracle):
    """src/minimizer.rs:726-742"""
    seq = xorshift_dna(5000)
    p = Params(k=19, w=10)
    anchors, tail = oracle.anchor_windows(p, seq, seq)
    assert anchors and not tail
    assert all(a.target_end == a.query_end and not a.adaptive_before for a in anchors)
    for x, y in zip(anchors, anchors[1:]):
        assert x.target_end + y.span <= y.target_end


def test_windows_split_cumulative_drift(oracle):
    """src/minimizer.rs:744-760"""
    ch = [oracle.Anchor(t, q, 19, False)
          for t, q in [(100, 100), (200, 200), (300, 220), (400, 230), (600, 600)]]
    windows, _ = oracle.partition_windows(ch, 800, 800, 100, 500)
    assert any(a.adaptive_before for a in windows)
    for x, y in zip(windows, windows[1:]):
        assert y.target_end >= x.target_end + y.span
        assert y.query_end >= x.query_end + y.span


def test_no_local_chain_without_minimum_score(oracle):
    """src/minimizer.rs:771-783"""
    anchors = [oracle.Anchor(100 + i * 20, 100 + i * 20, 19, False) for i in range(3)]
    assert oracle.chain(anchors, 19) == []


def test_chain_keeps_a_colinear_run_and_drops_an_off_diagonal_seed(oracle):
    """Chaining arithmetic on a hand-checked case (src/minimizer.rs:178-377):
    ten seeds 50 apart on the main diagonal score 19 + 9 x 19 = 190 >= 100 as
    one local chain. A seed 400 bases off the diagonal hangs off it with a
    step of 19 - min(0.19 x 400, log2 401) = 11, so its own local chain
    gains 11 < 100 and is dropped."""
    diag = [oracle.Anchor(100 + 50 * i, 100 + 50 * i, 19, False) for i in range(10)]
    stray = oracle.Anchor(330, 730, 19, False)
    got = oracle.chain(sorted(diag + [stray]), 19)
    assert got == diag


def linear(o, params, reference):
    """src/align.rs:1352-1358"""
    g = o.Graph(params)
    g.add_path(0, o.align(params, g, reference))
    g.toposort(False)
    return g


def test_seeded_windows_cover_the_query(oracle):
    """src/align.rs:1525-1559: the query is split into more than two windows
    and the seeded path has one record per query base plus the source."""
    p = Params(seeding=True, k=7, w=4, tr_dist=0, win_max=80, win_band=20)
    r = Lcg(123981)
    reference = bytes(b"ACGT"[r.next() % 4] for _ in range(2400))
    q = bytearray(reference)
    q[800:800] = b"TGCATGC"
    del q[1600:1605]
    q[2000] = ord("C") if q[2000] == ord("A") else ord("A")
    q = bytes(q)
    g = linear(oracle, p, reference)
    windows, _ = oracle.anchor_windows(p, q, reference)
    assert len(windows) > 2
    path = oracle.align(p, g, q)
    assert len(path) == len(q) + 1
    # Query bases in order: records that are not graph-only carry them.
    enc = {c: oracle.encode(p, c) for c in b"ACGTN"}
    assert [b.base for b in path[1:]] == [enc[c] for c in q]
    # Matched graph nodes strictly increase along the linear reference.
    nodes = [b.node for b in path[1:] if b.node is not None]
    assert nodes == sorted(set(nodes))


def test_seeded_workflow_renders_every_read(oracle):
    """tests/avx512.rs:145-180 (progressive=false; threads do not apply)."""
    reference = dna(1409)
    insertion = reference[:417] + dna(25) + reference[417:]
    deletion = reference[:803] + reference[820:]
    reads = records([reference, insertion, deletion])
    p = Params(seeding=True, k=7, w=3, tr_dist=0, win_max=80, win_band=24)
    g = oracle.align_sequences(p, reads)
    validate_msa(g.msa(p, reads), reads)
    validate_gfa(g.gfa(p, reads), reads)
    assert g.consensus(p)


def xorshift_prefix_records(n_reads=12, base=600):
    """src/minimizer.rs:753-762: random_dna(600 + i) restarts the xorshift
    stream each call, so the reads are prefixes of one sequence."""
    from minipoa_hip.params import Sequence
    return [Sequence(name=str(i), seq=xorshift_dna(base + i)) for i in range(n_reads)]


def guide_order_by_definition(oracle, p, seqs):
    """src/minimizer.rs:505-578 from the oracle's (pinned) sketch with plain
    dictionaries: multiset Jaccard x (0.51 len_i + 0.49 len_j) / max len,
    best ordered pair, then greedy by summed similarity, ties to lower index."""
    from collections import Counter, defaultdict
    n = len(seqs)
    if not p.progressive or n < 2 or len(p.alphabet) > 5:
        return list(range(n))
    groups = defaultdict(Counter)
    for rid, s in enumerate(seqs):
        for m in oracle.sketch(s.seq, p.k, p.w):
            groups[m.value][rid] += 1
    if not groups:
        return list(range(n))
    total = [0] * n
    shared = [[0] * n for _ in range(n)]
    for counts in groups.values():
        if len(counts) > n * p.mm_filter_ratio:
            continue
        for a, ca in counts.items():
            total[a] += ca
            for b, cb in counts.items():
                if b < a:
                    shared[a][b] += min(ca, cb)
                    shared[b][a] += min(ca, cb)
    max_len = float(max(1, max(len(s.seq) for s in seqs)))
    sim = [[0.0] * n for _ in range(n)]
    best = (-1.0, 0, 1)
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            c = shared[i][j]
            u = total[i] + total[j] - c
            jac = 0.0 if u == 0 else c / u * (0.51 * len(seqs[i].seq) / max_len
                                               + 0.49 * len(seqs[j].seq) / max_len)
            sim[i][j] = jac
            if jac > best[0]:
                best = (jac, i, j)
    order = [best[1], best[2]]
    used = set(order)
    sums = [sim[best[1]][v] + sim[best[2]][v] for v in range(n)]
    while len(order) < n:
        nxt = min((v for v in range(n) if v not in used), key=lambda v: (-sums[v], v))
        order.append(nxt)
        used.add(nxt)
        for v in range(n):
            if v not in used:
                sums[v] += sim[nxt][v]
    return order


def test_guide_order_is_permutation_and_deterministic(oracle):
    """src/minimizer.rs:753-776"""
    from minipoa_hip.params import Sequence
    seqs = xorshift_prefix_records()
    p = Params(progressive=True)
    order = oracle.guide_order(p, seqs)
    assert order == oracle.guide_order(p, seqs)
    assert sorted(order) == list(range(len(seqs)))
    short = [Sequence(seq=b"ACGT")] * 3
    assert oracle.guide_order(p, short) == [0, 1, 2]
    assert oracle.guide_order(Params(progressive=False), seqs) == list(range(12))


@pytest.mark.parametrize("seed", range(4))
def test_guide_order_matches_its_definition(oracle, seed):
    """Diverged copies of two templates in shuffled order: the oracle's
    grouping and greedy order equal the dictionary computation, and reads of
    one template come out next to each other."""
    import random as _r
    from minipoa_hip.params import Sequence
    rng = _r.Random(seed)
    t1, t2 = xorshift_dna(1500), bytes(rng.choice(b"ACGT") for _ in range(1300))

    def mut(t):
        return bytes(c if rng.random() > 0.05 else rng.choice(b"ACGT") for c in t)

    seqs = [Sequence(seq=mut(t1 if k % 2 else t2)) for k in range(9)]
    rng.shuffle(seqs)
    for p in (Params(progressive=True, k=15, w=5), Params(progressive=True, k=11, w=3, mm_filter_ratio=0.9),
              Params(progressive=True, mm_filter_ratio=0.0)):
        assert oracle.guide_order(p, seqs) == guide_order_by_definition(oracle, p, seqs)


def test_seeded_progressive_workflow_renders_every_read(oracle):
    """tests/avx512.rs:145-180 with progressive = true."""
    reference = dna(1409)
    insertion = reference[:417] + dna(25) + reference[417:]
    deletion = reference[:803] + reference[820:]
    reads = records([reference, insertion, deletion])
    p = Params(seeding=True, k=7, w=3, tr_dist=0, win_max=80, win_band=24, progressive=True)
    g = oracle.align_sequences(p, reads)
    validate_msa(g.msa(p, reads), reads)
    validate_gfa(g.gfa(p, reads), reads)


def test_seeded_batch_driver_is_thread_safe(oracle):
    """The job-parallel driver (the CPU baseline) gives every seeded job the
    same bytes and cells on one thread and on eight: the seeding's sorts keep
    no state shared between threads."""
    from minipoa_hip import synth
    jo, ro, b = synth.seeded(seed=20264, njobs=8, length=3000, nreads=8)
    p = Params(seeding=True)
    c1, h1 = oracle.batch_run(p, jo, ro, b, 1)
    c8, h8 = oracle.batch_run(p, jo, ro, b, 8)
    assert c1 == c8
    assert (h1 == h8).all() and (h1 != 0).all()
