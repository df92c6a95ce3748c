"""Shared test helpers: the reference's LCG streams, an independent textbook
Gotoh scorer, path re-scoring, and brute-force DAG optima."""
from __future__ import annotations

import itertools

MASK = (1 << 64) - 1
MATCH, DELETE, INSERT = 1, 2, 4
NEG = -(1 << 50)


class Lcg:
    """state = state * 6364136223846793005 + inc (wrapping); value = state >> 32
    (the generator every reference test uses, e.g. src/align.rs:1428-1434)."""

    def __init__(self, seed, inc=1):
        self.state = seed & MASK
        self.inc = inc

    def next(self):
        self.state = (self.state * 6364136223846793005 + self.inc) & MASK
        return self.state >> 32

    def i31(self):
        """`(state >> 32) as i32 & 0x7fffffff` (src/align.rs:1571-1574)."""
        return self.next() & 0x7FFFFFFF


def pairwise_optimum(params, reference, query):
    """Textbook global affine-gap (Gotoh) optimum, restated from the
    reference test helper src/align.rs:1359-1390."""
    width = len(query) + 1
    size = (len(reference) + 1) * width
    best = [NEG] * size
    dele = [NEG] * size
    ins = [NEG] * size
    best[0] = 0
    e = params.gap_ext
    o = params.gap_open + e
    for i in range(len(reference) + 1):
        for j in range(len(query) + 1):
            pos = i * width + j
            if i > 0:
                dele[pos] = max(best[pos - width] + o, dele[pos - width] + e)
            if j > 0:
                ins[pos] = max(best[pos - 1] + o, ins[pos - 1] + e)
            if i > 0 and j > 0:
                diag = best[pos - width - 1] + params.score(
                    params.encode(reference[i - 1]), params.encode(query[j - 1]))
            else:
                diag = best[pos]
            best[pos] = max(diag, dele[pos], ins[pos])
    return best[-1]


def alignment_score(params, node_rank_base, path, reference_len):
    """Re-score an alignment path (src/align.rs:1391-1425).
    node_rank_base(id) -> (rank, base)."""
    score = 0
    previous = 0
    last = [0]

    def add_gap(op):
        nonlocal score
        score += params.gap_ext
        if last[0] != op:
            score += params.gap_open
        last[0] = op

    for item in path[1:]:
        target = item.node if item.node is not None else item.aligned
        if target is not None:
            rank, base = node_rank_base(target)
            for _ in range(previous + 1, rank):
                add_gap(DELETE)
            score += params.score(base, item.base)
            last[0] = MATCH
            previous = rank
        else:
            add_gap(INSERT)
    for _ in range(previous, reference_len):
        add_gap(DELETE)
    return score


def dag_paths(succ, src, dst):
    """All source->sink paths of a small DAG (list of node-id lists)."""
    out = []
    stack = [(src, [src])]
    while stack:
        u, p = stack.pop()
        if u == dst:
            out.append(p)
            continue
        for v in succ[u]:
            stack.append((v, p + [v]))
    return out


def dag_dp_optimum(params, node_bases, succ, query):
    """Optimum of the sequence-to-DAG DP by brute force on a tiny graph.

    The DP aligns [N] + query + [N] against source .. sink, with the source
    pinned to column 0 (score(N, N)) and the sink scored as base N. That is
    score(N, N) plus the best textbook Gotoh optimum of (query + 'N') against
    the bases of any source->sink path after the source."""
    alpha = bytes(params.alphabet)
    nn = params.encode("N")
    best = None
    for p in dag_paths(succ, 0, 1):
        ref = bytes(alpha[node_bases[u]] for u in p[1:])
        s = pairwise_optimum(params, ref, bytes(query) + b"N")
        best = s if best is None else max(best, s)
    return params.score(nn, nn) + best


def dna(length, seed=0x5A71B301):
    """tests/avx512.rs:5-13"""
    r = Lcg(seed)
    return bytes(b"ACGT"[r.next() % 4] for _ in range(length))


def validate_msa(output, reads):
    """tests/avx512.rs:27-52"""
    lines = output.splitlines()
    rows = {}
    for k in range(0, len(lines), 2):
        rows[lines[k][1:].split()[0]] = lines[k + 1]
    assert len(rows) == len(reads)
    assert len({len(r) for r in rows.values()}) <= 1
    for read in reads:
        assert rows[read.name].replace("-", "").encode() == bytes(read.seq)


def validate_gfa(output, reads):
    """tests/avx512.rs:54-91"""
    nodes, links, paths = {}, set(), {}
    for line in output.splitlines():
        f = line.split("\t")
        if f[0] == "S":
            nodes[f[1]] = f[2]
        elif f[0] == "L":
            links.add((f[1], f[3]))
        elif f[0] == "P":
            paths[f[1]] = [] if f[2] == "*" else [x[:-1] for x in f[2].split(",")]
    assert len(paths) == len(reads)
    for read in reads:
        path = paths[read.name]
        assert "".join(nodes[i] for i in path).encode() == bytes(read.seq)
        assert all((a, b) in links for a, b in zip(path, path[1:]))
