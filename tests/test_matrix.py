"""`io::load_matrix` (the -m option) on the CPU: the reference's own cases
(tests/io_cli.rs:326-372) and its HOXD70 matrix file; a HOXD70-scored
alignment runs on the device against the oracle."""
from pathlib import Path

import pytest

import minipoa_hip as mh
from minipoa_hip import Params, records

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(tmp_path, text):
    f = tmp_path / "m.mtx"
    f.write_bytes(text)
    return mh.load_matrix(Params(), f)


def test_scoring_matrix_respects_labels_and_row_order(tmp_path):
    """tests/io_cli.rs:326-335"""
    p = load(tmp_path, b"# reordered alphabet\n N G A C T\nT 0 -3 -4 -5 7\nA 0 -2 6 -4 -5\n"
                       b"N 8 0 0 0 0\nC 0 -1 -2 9 -3\nG 0 10 -2 -3 -4\n")
    assert p.score(p.encode("a"), p.encode("A")) == 6
    assert p.score(p.encode("g"), p.encode("C")) == -3
    assert p.score(p.encode("U"), p.encode("T")) == 7
    assert p.encode("?") == p.encode("N")
    mh.validate(p)


def test_amino_acid_matrix_has_no_nucleotide_remapping(tmp_path):
    """tests/io_cli.rs:337-353"""
    alphabet = b"ACGTNU"
    text = "A C G T N U\n" + "".join(
        chr(r) + "".join(" 5" if r == c else " -4" for c in alphabet) + "\n" for r in alphabet)
    p = load(tmp_path, text.encode())
    assert p.encode("U") != p.encode("T")
    assert p.score(p.encode("U"), p.encode("U")) == 5


@pytest.mark.parametrize("text", [
    "# empty\n",
    "A A\nA 1 2\n",
    "A C\nA 1\nC 1 2\n",
    "A C\nA 1 2\n",
    "A C\nA 1 2\nA 2 1\n",
    "A C\nA 1 x\nC 1 2\n",
])
def test_rejects_malformed_matrix_instead_of_partial_initialization(tmp_path, text):
    """tests/io_cli.rs:355-372: an error, and params stay as they were."""
    p = Params()
    before = (bytes(p.alphabet), list(p.matrix))
    f = tmp_path / "bad.mtx"
    f.write_text(text)
    with pytest.raises(mh.HipPoaError) as e:
        mh.load_matrix(p, f)
    assert e.value.status == 11
    assert (bytes(p.alphabet), list(p.matrix)) == before


def test_hoxd70_matrix_file():
    """test-data/HOXD70.mtx as shipped with the reference."""
    p = mh.load_matrix(Params(), GOLDEN / "HOXD70.mtx")
    assert bytes(p.alphabet) == b"ACGTN"
    s = lambda a, b: p.score(p.encode(a), p.encode(b))
    assert (s("A", "A"), s("A", "C"), s("C", "G"), s("T", "G"), s("N", "N"), s("N", "A")) == \
        (91, -114, -125, -144, 100, 0)
    mh.validate(p)


@pytest.mark.gpu
def test_hoxd70_scored_alignment_matches_oracle(oracle):
    """Large per-base scores (the file asks for large gap penalties) through
    the device pipeline, against the oracle with the same loaded matrix."""
    from minipoa_hip import synth

    p = mh.load_matrix(Params(gap_open=-400, gap_ext=-30, adaptive=True), GOLDEN / "HOXD70.mtx")
    reads = synth.c1(seed=70, length=900, nreads=7)
    seqs = records(reads)
    ctx = mh.Context(0)
    try:
        g = mh.align_sequences(ctx, p, seqs)
        og = oracle.align_sequences(p, seqs)
        assert g.consensus(p) == og.consensus(p)
        assert g.msa(p, seqs) == og.msa(p, seqs)
        assert g.gfa(p, seqs) == og.gfa(p, seqs)
    finally:
        ctx.close()
