"""Known answers for the reference's ordering rules, each worked out by hand
from the reference's code, so that a plausible one-line change of the rule
(made in the oracle by tests/test_oracle_mutants.py) fails its case.

Conventions used in every derivation below (src/align.rs:587-1347):
  * rows are graph nodes in rank order, row 0 the source, the last row the
    sink (scored as base N); columns 0..m-1 with m = |query| + 2, column 0
    the source's N and column m-1 the sink's N;
  * o = gap_open + gap_ext, e = gap_ext; s(N,N) = match, s(N,x) = 0;
  * M[0][0] = s(N,N), M[0][j] = I[0][j] = M[0][j-1] + o or I[0][j-1] + e;
  * row i: Mdiag[j] = max_p M[p][j-1], D[j] = max_p max(M[p][j] + o,
    D[p][j] + e), I[j] = max(I[j-1] + e, M[j-1] + o), M[j] = max(Mdiag[j] +
    s(b_i, q_j), D[j], I[j]);
  * traceback from (n-1, m-1) with op = ALL: preferred diagonal (filtered
    by w >= sum(w)/10) -> delete (open, then extend) -> insert (open, then
    extend) -> diagonal; `choose` keeps the first of the heaviest
    predecessors that reproduce the score (src/align.rs:1201-1225).
Node ids: 0 source, 1 sink, then nodes in creation order from 2. Bases are
indices into ACGTN (A 0, C 1, G 2, T 3, N 4). A path is the AlignedBase
list `poa` returns: (node, base, aligned), None for absent.
"""
from minipoa_hip import AlignedBase, Params

N = 4
A, C, G, T = 0, 1, 2, 3


def ab(node, base, aligned=None):
    return AlignedBase(node, base, aligned)


class Case:
    """A graph (node bases in creation order; edges in insertion order, a
    repeated edge adds weight), a query and parameters, with the answer."""

    def __init__(self, name, params, bases, edges, query, path=None, cells=None, adaptive=False,
                 band=None, ranks=None, consensus=None):
        self.name = name
        self.params = params
        self.bases = bases
        self.edges = edges
        self.query = query
        self.path = path
        self.cells = cells
        self.adaptive = adaptive or params.adaptive
        self.band = params.b if band is None else band
        self.ranks = ranks
        self.consensus = consensus

    def build(self, graph_cls):
        g = graph_cls(self.params)
        for b in self.bases:
            g.add_node(b)
        for a, b in self.edges:
            g.add_edge(a, b)
        g.toposort(False)
        return g


def zero_open(**kw):
    return Params(gap_open=0, **kw).rebuild_matrix()


# ---- choose: heaviest predecessor, first of equals (src/align.rs:1201-1225)

CHOOSE_TIE = Case(
    "choose_equal_weights_first_incoming", Params(), [A, A],
    [(0, 2), (2, 1), (0, 3), (3, 1)], b"A",
    path=[ab(0, N), ab(2, A)])
CHOOSE_TIE.__doc__ = """Two parallel A nodes X1 (2) and X2 (3) between source
and sink, query "A". Ranks: 0, X2, X1, sink (LIFO). Both rows are equal:
M[X][1] = s(N,N) + s(A,A) = 4, and at the sink M[3][2] = 4 + s(N,N) = 6 from
either. The sink's preferred diagonal (N == N) has two candidates with
weight 1 each; sink.incoming = [X1, X2], so strict > keeps X1. At X1, j = 1,
the diagonal from the source. Path: source, X1 as A."""

CHOOSE_HEAVIER = Case(
    "choose_heavier_predecessor", Params(), [A, A],
    [(0, 2), (2, 1), (0, 3), (3, 1), (3, 1)], b"A",
    path=[ab(0, N), ab(3, A)])
CHOOSE_HEAVIER.__doc__ = """As CHOOSE_TIE with the edge X2 -> sink twice
(weight 2): the same scores, and the heavier X2 wins over the first X1."""


# ---- preferred-diagonal filter w >= sum(w) / 10 (src/align.rs:1227-1229)

def _filter_case(name, w_light, w_heavy, path):
    """Graph: L (C, 2) and T (A, 3) with source -> L -> T -> sink, and a
    second route source -> Z (C, 4) -> H (C, 5) -> T. Edge L -> T has weight
    w_light, H -> T weight w_heavy (w_light + w_heavy = 20, so the filter
    threshold at T is 2). gap_open 0, gap_ext -2 (o = e = -2), match 2,
    mismatch -4; query "AA" (m = 4).
    Ranks: 0, Z, H, L, T, sink. Rows (M by column 0..3):
      source  2  0 -2 -4
      Z, L    0 -2 -4 -2      (C: mismatch -4 against A)
      H      -2 -4 -6 -4
      T      -2  2  0 -2      I[T] = NEG -4 0 -2
    T at column 2: the diagonal is max(M[L][1], M[H][1]) + 2 = -2 + 2 = 0,
    from L only; D = -6; I = M[T][1] + o = 0. So M = I = diagonal = 0.
    Sink: M[5][3] = M[T][2] + s(N,N) = 2, its only predecessor T (weight 1,
    threshold 0).
    At (T, 2): A == q2 and the preferred diagonal is L with weight w_light.
      w_light = 1 < 2: filtered out. M != D; M == I and I == M[T][1] + o:
        q2 is an insertion, op = ALL (gap_open == 0). At (T, 1) the
        preferred L is filtered again, M != D, M != I, and the plain
        diagonal takes L: T is q1. At (L, 0) M == D: open from the source.
        Path: source, T as A (q1), insertion A (q2).
      w_light = 2 >= 2: kept, T is q2 and the walk goes to (L, 1): C != A,
        M == D = -2 == M[0][1] + o, so L is deleted (op = ALL); at (0, 1)
        M == I == M[0][0] + o: q1 is an insertion.
        Path: source, insertion A (q1), T as A (q2)."""
    return Case(name, zero_open(), [C, A, C, C],
                [(0, 2), (2, 3)] * 1 + [(2, 3)] * (w_light - 1) + [(3, 1), (0, 4), (4, 5)]
                + [(5, 3)] * w_heavy, b"AA", path=path)


FILTER_LIGHT = _filter_case("preferred_filtered_below_tenth", 1, 19,
                            [ab(0, N), ab(3, A), ab(None, A)])
FILTER_EDGE = _filter_case("preferred_kept_at_exactly_tenth", 2, 18,
                           [ab(0, N), ab(None, A), ab(3, A)])


# ---- delete: open before extend, and the gap_open == 0 state rule
# (src/align.rs:1237-1258)

OPEN_BEFORE_EXTEND = Case(
    "delete_open_before_extend", zero_open(), [A, A, C],
    [(0, 2), (2, 3), (3, 4), (4, 1)], b"A",
    path=[ab(0, N), ab(3, A)])
OPEN_BEFORE_EXTEND.__doc__ = """Chain source -> A1 (2) -> A2 (3) -> S (C, 4) ->
sink, query "A"; gap_open 0, gap_ext -2 (o = e = -2).
Rows (M / D by column 0..2):
  A1  M 0 4 2     D 0 -2 -4
  A2  M -2 2 4    D -2 2 0
  S   M -4 0 2    D -4 0 2
Sink: M[4][2] = M[S][1] + s(N,N) = 2 (preferred diagonal to S).
At (S, 1): C != A, M = D = 0, and both D == M[A2][1] + o = 0 and D ==
D[A2][1] + e = 0. Open comes first: to A2 with op = ALL (gap_open == 0).
At (A2, 1): A == q1 and M = 2 = M[A1][0] + 2: the preferred diagonal, A2
takes the A. (Extending first would enter A2 with op = DELETE, skip the
diagonal and give the A to A1.) Path: source, A2 as A."""


# ---- insert: open before extend (src/align.rs:1260-1288)

INSERT_OPEN_BEFORE_EXTEND = Case(
    "insert_open_before_extend", zero_open(), [A],
    [(0, 2), (2, 1)], b"AAC",
    path=[ab(0, N), ab(None, A), ab(2, A), ab(None, C)])
INSERT_OPEN_BEFORE_EXTEND.__doc__ = """One node T (A, 2) between source and sink,
query "AAC" (m = 5); gap_open 0, gap_ext -2 (o = e = -2).
Row T (M / I by column 0..4): M 0 4 2 0 -2, I NEG -2 2 0 -2.
Sink: M[2][4] = M[T][3] + s(N,N) = 2 (preferred diagonal to T at column 3).
At (T, 3): A != C, M != D, M == I = 0 with I == M[T][2] + o = 0 and I ==
I[T][2] + e = 0. Open comes first: insertion of C, op = ALL. At (T, 2):
A == q2 and M = 2 = M[0][1] + 2: the diagonal, T takes the second A. At
(0, 1) the source inserts the first A. (Extending first would stay in
insertion mode at (T, 2), insert the second A and give T the first.)
Path: source, insertion A, T as A, insertion C."""



# ---- the diagonal from a predecessor whose band ends one column before
# (src/align.rs:764-794 fill overlap, :1201-1225 `choose` over in-band j - 1)

BAND_EDGE_DIAGONAL = Case(
    "diagonal_from_band_ending_one_column_before",
    Params(gap_open=-2, gap_ext=-2, match_score=2, mismatch=-1, b=0, f=1000).rebuild_matrix(),
    [T, G], [(0, 2), (0, 3), (2, 1), (2, 1), (3, 1)], b"TCTGATA",
    path=[ab(0, N), ab(None, T), ab(None, C), ab(None, T), ab(3, G), ab(None, A), ab(None, T),
          ab(None, A)], cells=40)
BAND_EDGE_DIAGONAL.__doc__ = """Source -> X (T, 2) -> sink (weight 2) and source ->
Y (G, 3) -> sink (weight 1); query "TCTGATA" (m = 9); gap_open -2, gap_ext
-2 (o = -4, e = -2), match 2, mismatch -1; b = 0 (f = 1000 adds nothing).
Bands (head, m - tail): source (0, 6) and X, Y (1, 7) give [0, 8); the sink
(2, 8) gives [0, 16). So the sink's column 8 is in no predecessor's band:
its fill has no diagonal and no deletion there.
Rows (M by column 0..7):
  source  2 -2 -4 -6  -8 -10 -12 -14
  X      -2  4  0 -2  -4  -6  -8 -10
  Y      -2  1 -3 -5  -4  -8 -10 -12   (I[Y][5..7] = -8 -10 -12)
  sink   -4  0  4  0  -2  -4  -6  -8, and M[sink][8] = I = -8 - 2 = -10.
At (sink, 8) op = ALL: N == N and `choose` looks at column 7 of every
predecessor whose band holds column 7, which both do: X gives -10 + 2 = -8,
Y gives -12 + 2 = -10 = M. So Y is the diagonal (weight 1 >= 3/10), though
the fill never took that diagonal. At (Y, 7) M == I == I[Y][6] + e: A, T are
insertions (extend), and at (Y, 5) I == M[Y][4] + o: A is an insertion
(open, op = MATCH|DELETE). At (Y, 4) G == G and M = -4 = M[0][3] + 2: Y
takes the G. The source inserts T, C, T. (A `choose` that also required
column 8 in the predecessor's band would find no diagonal at (sink, 8),
insert there and give the first T to X.)"""


# ---- head/tail: heaviest edge, ties to larger support (src/align.rs:623-678)

HEAD_TAIL_SUPPORT = Case(
    "head_and_tail_support_ties",
    Params(b=0, f=1000),
    [A] * 9 + [A, A],
    [(0, 2)] + [(k, k + 1) for k in range(2, 10)] + [(10, 12), (0, 11), (11, 12), (12, 1)],
    b"A" * 10,
    path=[ab(0, N), ab(11, A)] + [ab(None, A)] * 8 + [ab(12, A)],
    cells=112)
HEAD_TAIL_SUPPORT.__doc__ = """Long route source -> Q1..Q9 (2..10) -> J (12) and
short route source -> P1 (11) -> J, then J -> sink; all A, every edge weight
1; query A x 10 (m = 12); b = 0, f = 1000, so band = 0.
Ranks: 0, P1, Q1..Q9, J, sink (rows 0..12).
Head pass: Q_k head k, support k; P1 head 1, support 1. J's in-edges tie on
weight (J.incoming = [Q9, P1]); the larger support is Q9's (9 > 1), so head
J = 10, support 10; sink head 11, support 11.
Tail pass (support shared): sink 1; J 2 (support 12); Q_k tail 12 - k
(support 22 - k); P1 tail 3 (support 13); the source ties on weight between
Q1 (support 21) and P1 (13) and takes Q1: tail 12.
Static band 0: rows cover the 8-blocks from min(h, m-t) to max(h, m-t):
source h 0, m-t 0 -> [0,8) 8; P1 1..9 -> [0,16) 16; Q_k k..k -> 8 each (72);
J 10..10 -> [8,16) 8; sink 11..11 -> 8. Cells 8+16+72+8+8 = 112.
(A smaller-support head tie gives J head 2 and the sink 3: rows [0,16), 128
cells; a smaller-support tail tie gives the source tail 4: [0,16), 120.)
Predecessor transitions only cover the overlap of the two rows' bands
(src/align.rs:751-797): Q8's row [8,16) and Q7's [0,8) do not overlap, so
Q8 and Q9 receive nothing and the long route is cut. J ([8,16)) is reached
only from P1 ([0,16)): P1 takes q1 (M = 4), q2..q9 are insertions (o + 7e
= -20) and J takes q10; the sink adds s(N,N): score -12.
Path: source, P1 as A, eight insertions of A, J as A."""


# ---- adaptive band: streak >= 30 recentres (src/align.rs:845-865)

ADAPTIVE_RECENTRE = Case(
    "adaptive_streak_30_recentres",
    Params(b=2, f=10000, adaptive=True),
    [C] * 16 + [A] * 20,
    [(0, 2)] + [(k, k + 1) for k in range(2, 17)] + [(17, 18)]
    + [(k, k + 1) for k in range(18, 37)] + [(37, 1)] + [(0, 18)] * 5,
    b"C" * 16 + b"A" * 20,
    path=[ab(0, N)] + [ab(k, C) for k in range(2, 18)] + [ab(k, A) for k in range(18, 38)],
    cells=680)
ADAPTIVE_RECENTRE.__doc__ = """Chain source -> B1..B16 (C, 2..17) -> A1..A20 (A,
18..37) -> sink, plus a heavy shortcut source -> A1 (weight 5); query C^16
A^20 (m = 38); b = 2, f = 10000 (band 2); default scores.
Heads: B_k k; A_1 1 (the shortcut outweighs B16), A_i i. Tails: A_i 22 - i,
B_k 38 - k, source 22 (its heaviest out-edge is the shortcut). So m - t is
k on B_k and 16 + i on A_i: the alignment runs 16 columns right of head.
The source row pushes nothing (its update is skipped). Peaks (first
argmax of M): B_k at k, A_i at 16 + i (all matches); every row matches
(base == q[peak], a predecessor's peak + 1 == peak): streak B_k = k, A_i =
16 + i. A_14 reaches 30 and recentres: offsets = (16 + 14) - 14 = 16, pushed
to A_15; every later A row does the same; extra stays 0.
Rows, band [a - 2, b + 2] with a = min(peak_left, h + off_l, m - t), b =
max(peak_right, h + off_r, m - t):
  source a 0, b 16                       -> [0,24)       24
  B_k    a k-1, b k                      -> 8,8,8,8,8,16,16,16,16,16,8,8,8,
                                            16,16,16    192
  A_1    peak_left 16 (B16 only), a 1, b 17 -> [0,24)   24
  A_2..A_14 a i, b 16+i                  -> 24x4, 32x4, 24x4, 32   352
  A_15..A_20 a 15+i, b 16+i (recentred)  -> 16,16,16,16,8,8   80
  sink   a 36, b 37                      -> [32,40)       8
Cells 680. (Without recentring A_15..A_20 and the sink keep a = i: 784;
recentring from 31 only widens A_15: 696.)"""


# ---- adaptive band: extra grows by max_streak + 1 up to 5 x band (:845-850)

ADAPTIVE_EXTRA = Case(
    "adaptive_extra_growth_and_cap",
    Params(b=1, f=10000, adaptive=True, mismatch=-1, gap_open=-4, gap_ext=-4).rebuild_matrix(),
    [A] * 24,
    [(0, 2)] + [(k, k + 1) for k in range(2, 25)] + [(25, 1)],
    b"C" * 24,
    path=[ab(0, N)] + [ab(None, C, k) for k in range(2, 26)],
    cells=464)
ADAPTIVE_EXTRA.__doc__ = """Chain source -> X1..X24 (A, 2..25) -> sink, query C^24
(m = 26); b = 1, f = 10000 (band 1, cap 5); match 2, mismatch -1, gap_open
-4, gap_ext -4 (o = -8, e = -4), so the all-mismatch diagonal is the best
in every row: peak X_k = k, M = 2 - k. No row matches (A != C), so
max_streak is 0 and every row adds 1 to extra, capped at 5; X_k's band
uses its predecessor's value, min(k - 1, 5). Heads k, m - t = k; band
[k - 1 - 1 - extra, k + 1 + extra] (X_1: a = 1):
  source [0,8) 8; X1..X3 8 each; X4..X9 16 each; X10..X14 24 each;
  X15..X17 16 each; X18..X22 24 each; X23, X24 16 each; sink (a 24, b 25,
  extra 5) [16,32) 16.
Cells 8+24+96+120+48+120+32+16 = 464. The path is the diagonal of
mismatches: each X_k aligned with C (no C node in its column)."""


# ---- graph: LIFO toposort and consensus ties (src/graph.rs:124-237)

TOPOSORT_LIFO = Case(
    "toposort_lifo", Params(), [A, C], [(0, 2), (0, 3), (2, 1), (3, 1)], b"",
    ranks=[0, 3, 2, 1])
TOPOSORT_LIFO.__doc__ = """Source -> X1 (2), X2 (3) -> sink. The stack starts
with [0]; popping 0 pushes X1 then X2 (out-edge order); X2 is popped first
(rank 1), then X1 (rank 2) frees the sink (rank 3): order 0, 3, 2, 1."""

CONSENSUS_INDEX_TIE = Case(
    "consensus_lowest_index_tie", Params(), [A, C], [(0, 2), (2, 1), (0, 3), (3, 1)], b"",
    consensus=b"A")
CONSENSUS_INDEX_TIE.__doc__ = """Source -> X1 (A) -> sink and source -> X2 (C) ->
sink, all weights 1: at the sink both in-edges weigh 1 and both
predecessors score 1, so the lower incoming index (X1) wins: "A"."""

CONSENSUS_SCORE_TIE = Case(
    "consensus_score_breaks_weight_tie", Params(), [C, G, T],
    [(0, 2), (2, 1), (0, 3), (3, 4), (4, 1)], b"",
    consensus=b"GT")
CONSENSUS_SCORE_TIE.__doc__ = """Source -> Z (C, 2) -> sink and source -> Y1 (G, 3)
-> Y2 (T, 4) -> sink, weights 1; sink.incoming = [Z, Y2]. Equal weights;
score Z = 1, score Y2 = 2: the larger score wins over the lower index:
"GT"."""

PATH_CASES = [CHOOSE_TIE, CHOOSE_HEAVIER, FILTER_LIGHT, FILTER_EDGE, OPEN_BEFORE_EXTEND,
              INSERT_OPEN_BEFORE_EXTEND, BAND_EDGE_DIAGONAL, HEAD_TAIL_SUPPORT, ADAPTIVE_RECENTRE,
              ADAPTIVE_EXTRA]
GRAPH_CASES = [TOPOSORT_LIFO, CONSENSUS_INDEX_TIE, CONSENSUS_SCORE_TIE]
