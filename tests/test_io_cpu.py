"""Input formats (the reference's io module; src/io.rs is not in its tree,
so these follow its tests and README): FASTA/FASTQ records, gzip input and
the GFA graph of the -i option. Known answers from tests/io_cli.rs:34-87,
374-419, 563-610; plus an export -> import invariant on graphs built from
reads: every base, edge weight and path comes back."""
import gzip
import random

import pytest

import minipoa_hip as mh
from helpers import dna
from minipoa_hip import AlignedBase, Params, Sequence


def test_fasta_multiline_crlf_comments_and_empty_records():
    """tests/io_cli.rs:34-46"""
    r = mh.parse_sequences(b"\n>alpha annotation words\r\nAC GT\r\nN\r\n>empty\r\n>last\r\nU\r\n")
    assert len(r) == 3
    assert (r[0].name, r[0].comment, r[0].seq) == ("alpha", "annotation words", b"ACGTN")
    assert r[1].seq == b""
    assert r[2].seq == b"U"


def test_fastq_multiline_quality_may_start_with_header_marker():
    """tests/io_cli.rs:48-56"""
    r = mh.parse_sequences(b"@alpha words\nAC\nGT\n+alpha\n@@\n++\n@beta\nNN\n+\nII\n")
    assert len(r) == 2
    assert r[0].seq == b"ACGT" and r[0].qual == b"@@++"
    assert r[1].qual == b"II"


@pytest.mark.parametrize("text", [b"ACGT\n", b">\nA\n", b"@a\nAC\n+\nI\n", b"@a\nA\n+\nIII\n",
                                  b"@a\nAA\n@b\nA\n+\nI\n", b">a\n@b\n"])
def test_malformed_sequence_files_are_errors(text):
    """tests/io_cli.rs:58-73"""
    with pytest.raises(mh.HipPoaError):
        mh.parse_sequences(text)


def test_empty_sequence_input_has_no_records():
    assert mh.parse_sequences(b"") == []


def test_concatenated_gzip_members_and_magic_detection(tmp_path):
    """tests/io_cli.rs:75-87: two gzip members in one file, detected by their
    magic (not the file name); a truncated stream is an error."""
    data = gzip.compress(b">a\nAC\n") + gzip.compress(b">b\nGT\n")
    f = tmp_path / "reads.data"
    f.write_bytes(data)
    r = mh.read_sequences(f)
    assert len(r) == 2 and r[1].seq == b"GT"
    with pytest.raises(mh.HipPoaError):
        mh.parse_sequences(data[:-3])
    # Plain text that happens to be valid is read as is.
    assert mh.parse_sequences(b">x\nACGT\n")[0].seq == b"ACGT"


def test_gfa_paths_include_single_base_segments_and_roundtrip():
    """tests/io_cli.rs:374-390"""
    p = Params()
    g, rec = mh.read_gfa(p, "H\tVN:Z:1.0\nS\tone\tA\nS\ttwo\tCG\nL\tone\t+\ttwo\t+\t0M\n"
                            "P\tread\tone+,two+\t*\n")
    assert rec[0].seq == b"ACG"
    assert len(g.paths[0][1]) == 3
    assert g.consensus(p)[0] == b"ACG"
    out = g.gfa(p, rec)
    again, seqs = mh.read_gfa(p, out)
    assert again.consensus(p)[0] == b"ACG"
    assert seqs[0].seq == b"ACG"


def test_gfa_links_without_paths_still_initialize_a_graph():
    """tests/io_cli.rs:392-399"""
    p = Params()
    g, rec = mh.read_gfa(p, "S\ta\tAC\nS\tb\tGT\nL\ta\t+\tb\t+\t0M\n")
    assert rec == []
    assert g.consensus(p)[0] == b"ACGT"


@pytest.mark.parametrize("text", [
    "S\ta\tA\nS\ta\tC\n",
    "S\ta\tA\nP\tx\tb+\t*\n",
    "S\ta\tA\nP\tx\ta-\t*\n",
    "S\ta\tA\nL\ta\t+\ta\t+\t0M\n",
    "S\ta\tA\nS\tb\tC\nL\ta\t+\tb\t+\t0M\nL\tb\t+\ta\t+\t0M\n",
    "S\ta\tA\nL\ta\t+\ta\t-\t0M\n",
    "S\ta\tA\nL\ta\t+\ta\t+\t1M\n",
    "S\ta\t*\n",
])
def test_rejects_invalid_gfa_references_cycles_or_unsupported_orientation(text):
    """tests/io_cli.rs:401-419"""
    with pytest.raises(mh.HipPoaError):
        mh.read_gfa(Params(), text)


@pytest.mark.parametrize("text", ["H\tVN:Z:1.0\nP\tempty\t*\t*\n",
                                  "H\tVN:Z:1.0\nS\ta\tAC\nP\tempty\t*\t*\nP\tnonempty\ta+\t*\n"])
def test_gfa_empty_paths_export_and_reimport(text):
    """tests/io_cli.rs:584-603"""
    p = Params()
    g, rec = mh.read_gfa(p, text)
    assert rec[0].name == "empty" and rec[0].seq == b""
    assert g.paths[0][1] == []
    out = g.gfa(p, rec)
    assert "P\tempty\t*\t*\n" in out
    loaded, again = mh.read_gfa(p, out)
    assert [(s.name, s.seq) for s in again] == [(s.name, s.seq) for s in rec]
    assert loaded.consensus(p)[0] == g.consensus(p)[0]


def test_rejects_gfa_paths_revisiting_single_base_segment():
    """tests/io_cli.rs:605-610"""
    with pytest.raises(mh.HipPoaError, match="revisits"):
        mh.read_gfa(Params(), "S\ta\tA\nP\tcycle\ta+,a+\t*\n")


def test_unsupported_gfa_records_are_errors_and_comments_are_skipped():
    """README.md: other record kinds produce explicit errors."""
    with pytest.raises(mh.HipPoaError, match="unsupported"):
        mh.read_gfa(Params(), "S\ta\tA\nC\ta\t+\tb\t+\t0\t0M\n")
    g, rec = mh.read_gfa(Params(), "# note\nS\ta\tACG\n\nP\tr\ta+\t*\n")
    assert rec[0].seq == b"ACG"


def _read_built_graph(oracle, params, reads):
    """A library graph holding `reads`, fused from the oracle's paths (the
    host fusion the device path uses, on the CPU)."""
    og = oracle.Graph(params)
    lg = mh.Graph(params)
    for sid, r in enumerate(reads):
        path = oracle.align(params, og, r)
        og.add_path(sid, path)
        og.toposort(False)
        lg.add_path(sid, path)
        lg.toposort(False)
    return lg


def test_exported_graph_reimports_every_base_edge_weight_and_path(oracle):
    """-r2 output of a graph built from reads, read back with read_gfa: the
    segments are the nodes (named id - 1), and every node's base, every edge
    with its weight (source and sink edges included) and every path come back
    under that renaming. Only the substitution-column groups are not in GFA."""
    rng = random.Random(11)
    p = Params()
    template = dna(300, seed=5)
    reads = []
    for _ in range(6):
        r = bytearray()
        for c in template:
            x = rng.random()
            if x < 0.05:
                continue
            r.append(rng.choice(b"ACGT") if x < 0.12 else c)
            if rng.random() < 0.03:
                r.append(rng.choice(b"ACGT"))
        reads.append(bytes(r))
    lg = _read_built_graph(oracle, p, reads)
    seqs = [Sequence(name=f"r{k}", seq=s) for k, s in enumerate(reads)]
    text = lg.gfa(p, seqs)
    g2, rec = mh.read_gfa(p, text)
    assert [(s.name, s.seq) for s in rec] == [(s.name, s.seq) for s in seqs]
    names = [line.split("\t")[1] for line in text.splitlines() if line.startswith("S\t")]
    new_of = {0: 0, 1: 1}
    for k, nm in enumerate(names):
        new_of[int(nm) + 1] = k + 2
    assert g2.n_nodes == lg.n_nodes
    for old, new in new_of.items():
        a, b = lg.node(old), g2.node(new)
        assert a["base"] == b["base"]
        assert sorted(zip([new_of[x] for x in a["incoming"]], a["in_weights"])) == \
            sorted(zip(b["incoming"], b["in_weights"])), old
        assert sorted(zip([new_of[x] for x in a["outgoing"]], a["out_weights"])) == \
            sorted(zip(b["outgoing"], b["out_weights"])), old
    assert [(s, [new_of[x] for x in ids]) for s, ids in lg.paths] == g2.paths
    assert g2.consensus(p)[0] == lg.consensus(p)[0]


def test_oracle_copy_of_a_loaded_graph_is_the_same_graph(oracle):
    """The oracle continues a GFA-loaded graph from a field-by-field copy:
    the copy renders the same GFA and MSA as the library graph."""
    p = Params()
    text = ("H\tVN:Z:1.0\nS\t1\tACGTTG\nS\t2\tA\nS\t3\tC\nS\t4\tGGT\nL\t1\t+\t2\t+\t0M\n"
            "L\t1\t+\t3\t+\t0M\nL\t2\t+\t4\t+\t0M\nL\t3\t+\t4\t+\t0M\n"
            "P\tx\t1+,2+,4+\t*\nP\ty\t1+,3+,4+\t*\nP\tz\t1+,2+,4+\t*\n")
    g, rec = mh.read_gfa(p, text)
    og = oracle_copy(oracle, p, g)
    assert og.gfa(p, rec) == g.gfa(p, rec)
    assert og.msa(p, rec) == g.msa(p, rec)
    assert og.consensus(p) == g.consensus(p)


def oracle_copy(oracle, params, g):
    nodes, aligned, order = g.snapshot()
    return oracle.graph_from_arrays(params, nodes, aligned, order, g.paths)
