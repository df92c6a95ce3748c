"""Mutation tests of the oracle: each ordering rule of the reference that
the oracle restates is changed, one at a time, by a plausible one-line
mistake (poa_oracle.c built with -DORACLE_MUTANT=k into a scratch
directory), and the hand-derived known answer for that rule must then
fail. The unmutated oracle passes all of them (tests/test_rules_kat.py)."""
import json
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

# mutant -> (the rule and its mistake, the known answers that must fail)
MUTANTS = {
    1: ("choose keeps the last of equally heavy predecessors (>= for >)", ["case:choose_equal_weights_first_incoming"]),
    2: ("choose ignores edge weights (first match wins)", ["case:choose_heavier_predecessor"]),
    3: ("no sum(w)/10 filter on the preferred diagonal", ["case:preferred_filtered_below_tenth"]),
    4: ("filter w > sum(w)/10 instead of >=", ["case:preferred_kept_at_exactly_tenth"]),
    5: ("delete: extend tried before open", ["case:delete_open_before_extend"]),
    6: ("delete open with gap_open == 0 leaves MATCH|INSERT", ["test:test_oracle::test_zero_gap_open_traceback_regression"]),
    7: ("insert open with gap_open == 0 leaves MATCH|DELETE",
        ["test:test_oracle::test_affine_alignment_handles_zero_gaps_and_ambiguous_bases"]),
    8: ("insert: extend tried before open", ["case:insert_open_before_extend"]),
    9: ("head tie goes to the smaller support", ["case:head_and_tail_support_ties"]),
    10: ("tail tie goes to the smaller support", ["case:head_and_tail_support_ties"]),
    11: ("adaptive band never recentres", ["case:adaptive_streak_30_recentres"]),
    12: ("adaptive band recentres from streak 31", ["case:adaptive_streak_30_recentres"]),
    13: ("extra grows by max_streak (no + 1)", ["case:adaptive_extra_growth_and_cap"]),
    14: ("extra capped at 4 x band", ["case:adaptive_extra_growth_and_cap"]),
    16: ("toposort pops first in, first out", ["case:toposort_lifo"]),
    17: ("consensus full tie keeps the highest incoming index", ["case:consensus_lowest_index_tie"]),
    18: ("consensus weight tie keeps the smaller path score", ["case:consensus_score_breaks_weight_tie"]),
    20: ("choose skips a diagonal predecessor whose band ends at j - 1",
         ["case:diagonal_from_band_ending_one_column_before"]),
}

RUNNER = r"""
import json, sys
sys.path[:0] = %(paths)r
import oracle, rule_kats as K
out = {}
for spec in %(specs)r:
    kind, name = spec.split(":", 1)
    try:
        if kind == "case":
            case = {c.name: c for c in K.PATH_CASES + K.GRAPH_CASES}[name]
            if case.path is not None:
                for kern in (oracle.poa_scalar_i64, oracle.poa_simd_i32):
                    g = case.build(oracle.Graph)
                    oracle.reset_cells()
                    r = oracle.result_or_error(kern, case.params, g, 0, 1, case.query,
                                               case.adaptive, case.band)
                    assert r == ("ok", case.path)
                    if case.cells is not None:
                        assert oracle.cells() == case.cells
            else:
                g = case.build(oracle.Graph)
                if case.ranks is not None:
                    assert g.rank == case.ranks
                if case.consensus is not None:
                    assert g.consensus(case.params)[0] == case.consensus
        else:
            mod, fn = name.split("::")
            getattr(__import__(mod), fn)(oracle)
        out[spec] = "pass"
    except Exception as e:
        out[spec] = "fail: " + type(e).__name__
print(json.dumps(out))
"""


def build(k, outdir):
    so = outdir / f"liboracle_m{k}.so"
    subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-fPIC", "-std=c11", "-w",
                    f"-DORACLE_MUTANT={k}", "-shared", "-pthread", "-o", str(so),
                    str(ROOT / "oracle" / "poa_oracle.c"), "-lm"], check=True)
    return so


def run(so, specs):
    code = RUNNER % {"paths": [str(ROOT / "tests"), str(ROOT / "minipoa-rs_amd"),
                               str(ROOT / "oracle")], "specs": specs}
    env = dict(os.environ, ORACLE_LIB=str(so))
    p = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300, check=True)
    return json.loads(p.stdout.strip().splitlines()[-1])


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        built = dict(zip(MUTANTS, ex.map(lambda k: build(k, d), MUTANTS)))
        results = dict(zip(MUTANTS, ex.map(lambda k: run(built[k], MUTANTS[k][1]), MUTANTS)))
    return results


@pytest.mark.parametrize("k", sorted(MUTANTS), ids=lambda k: f"m{k}")
def test_each_rule_mutant_fails_its_known_answer(mutants, k):
    rule, specs = MUTANTS[k]
    res = mutants[k]
    assert all(res[s].startswith("fail") for s in specs), (rule, res)


def test_unmutated_oracle_passes_the_same_known_answers():
    specs = sorted({s for _, ss in MUTANTS.values() for s in ss})
    res = run(ROOT / "oracle" / "liboracle_poa.so", specs)
    assert all(v == "pass" for v in res.values()), res
