"""Hand-derived known answers for the reference's ordering rules (the
derivations are in tests/rule_kats.py) on the oracle, with both of its
kernels, and on the library's host graph (toposort, consensus). The device
runs the same cases in tests/test_gpu_rules_kat.py; tests/test_oracle_mutants.py
shows that each case fails when its rule is changed."""
import pytest

import minipoa_hip as mh
import rule_kats as K


def run_oracle_case(oracle, case, kernel):
    """(path or error, cells) of one case on one oracle kernel."""
    g = case.build(oracle.Graph)
    oracle.reset_cells()
    r = oracle.result_or_error(kernel, case.params, g, 0, 1, case.query, case.adaptive, case.band)
    return r, oracle.cells()


@pytest.mark.parametrize("case", K.PATH_CASES, ids=lambda c: c.name)
@pytest.mark.parametrize("kernel", ["poa_scalar_i64", "poa_simd_i32"])
def test_oracle_path_known_answers(oracle, case, kernel):
    r, cells = run_oracle_case(oracle, case, getattr(oracle, kernel))
    assert r == ("ok", case.path)
    if case.cells is not None:
        assert cells == case.cells


@pytest.mark.parametrize("case", K.GRAPH_CASES, ids=lambda c: c.name)
def test_graph_known_answers_oracle_and_library(oracle, case):
    for cls in (oracle.Graph, mh.Graph):
        g = case.build(cls)
        if case.ranks is not None:
            assert g.rank == case.ranks, cls
        if case.consensus is not None:
            assert g.consensus(case.params)[0] == case.consensus, cls
