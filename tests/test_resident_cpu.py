"""The device-resident step's graph and planning code (resident.hpp) run on
one host thread beside the library's host graph and planner: after every
read, nodes, adjacency order, ranks and paths must be equal, and so must
the plans of every read's DP (banded and the full-DP retry). The host side
is itself held to the oracle (test_capi_cpu.py, test_oracle.py); the paths
applied here are the oracle's alignments of the reads."""
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, synth


def oracle_paths(oracle, params, reads):
    g = oracle.Graph(params)
    out = []
    for k, r in enumerate(reads):
        path = oracle.align(params, g, r)
        out.append(path)
        g.add_path(k, path)
        g.toposort()
    return out


@pytest.mark.parametrize("lds", [0, 65536])
@pytest.mark.parametrize("params", [Params(adaptive=True), Params(), Params(f=0), Params(b=3, adaptive=True)],
                         ids=["adaptive", "static", "full", "narrow"])
def test_resident_step_matches_host_graph_and_plans(oracle, params, lds):
    jo, ro, b = synth.jobs(seed=101, njobs=16, lmin=30, lmax=500, rmin=2, rmax=9)
    for j in range(len(jo) - 1):
        reads = synth.job_reads(jo, ro, b, j)
        assert mh.resident_check(params, reads, oracle_paths(oracle, params, reads), lds) == ""


def test_resident_step_on_batched_workload_jobs(oracle):
    """Two jobs of the 4,096-job workload at full length (L up to 2,000,
    up to 32 reads, adaptive band): graphs of thousands of nodes, staged in
    LDS and in global memory."""
    jo, ro, b = synth.jobs(seed=20262, njobs=6)
    p = Params(adaptive=True)
    for j in (0, 5):
        reads = synth.job_reads(jo, ro, b, j)
        paths = oracle_paths(oracle, p, reads)
        for lds in (0, 65536):
            assert mh.resident_check(p, reads, paths, lds) == ""


def test_resident_step_protein_full_dp(oracle):
    """Twenty-letter alphabet with a drawn matrix, -f0 (the protein workload)."""
    p, jo, ro, b = synth.protein_families(seed=7, nfam=4, lmin=40, lmax=200, members=6)
    assert p.f == 0 and len(p.alphabet) == 20
    for j in range(len(jo) - 1):
        reads = synth.job_reads(jo, ro, b, j)
        assert mh.resident_check(p, reads, oracle_paths(oracle, p, reads), 65536) == ""


def test_resident_step_leaves_device_where_host_needs_i64(oracle):
    """Scores past the i32 guard (src/align.rs:493-529): the host plans the
    i64 kernel exactly where the resident step hands the job back."""
    p = Params(match_score=1 << 16, mismatch=-(1 << 16), gap_open=-(1 << 16),
               gap_ext=-(1 << 15)).rebuild_matrix()
    reads = [b"ACGTTGCA" * k for k in (1, 3, 20, 40)]
    assert mh.resident_check(p, reads, oracle_paths(oracle, p, reads), 0) == ""


def test_resident_step_serial_form_on_repeated_nodes(oracle):
    """Paths whose elements resolve to the same node twice (a new member of
    one aligned column requested twice, an existing node named twice): the
    team form falls back to the reference's serial loop, with the host's
    outcome (src/graph.rs:82-122: the second request reuses the first's node,
    the repeated edge is a self edge and is ignored)."""
    p = Params()
    reads = [b"ACGT", b"ACGT", b"ACGT", b"ACGT"]
    first = oracle_paths(oracle, p, reads[:1])[0]
    # after read 0: source 0, sink 1, nodes 2..5 = A C G T (parents themselves)
    twice_new = [(0, 4, None), (None, 2, 2), (None, 2, 2), (3, 1, None), (4, 2, None), (5, 3, None)]
    twice_old = [(0, 4, None), (2, 0, None), (2, 0, None), (3, 1, None), (4, 2, None), (5, 3, None)]
    plain = [(0, 4, None), (2, 0, None), (None, 3, 3), (4, 2, None), (5, 3, None)]
    assert mh.resident_check(p, reads, [first, twice_new, twice_old, plain], 65536) == ""
    assert mh.resident_check(p, reads, [first, twice_new, twice_old, plain], 0) == ""


def test_resident_step_detects_the_same_cycle_as_the_host(oracle):
    """A path that closes a cycle: both sides report it on the same read."""
    p = Params()
    reads = [b"ACGT", b"ACGT"]
    first = oracle_paths(oracle, p, reads[:1])[0]
    loop = [(0, 4, None), (5, 3, None), (2, 0, None)]  # T -> A closes A->C->G->T->A
    assert mh.resident_check(p, reads, [first, loop], 0) == ""
