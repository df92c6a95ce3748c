"""The autotuner measures launch shapes on the device and writes a table the
library follows (MPOA_TUNING); results are the same with and without it,
since every shape computes the same DP."""
import os

import numpy as np
import pytest

import minipoa_hip as mh
from minipoa_hip import Params, synth, tune

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mh.Context(0)
    yield c
    c.close()


def test_tuned_table_is_written_and_followed_with_identical_results(ctx, tmp_path):
    path = tmp_path / "tuning.txt"
    chosen = tune.tune(ctx, str(path), buckets=[384, 1024], shapes=[(8, 1), (4, 4)], rings=[16],
                       cells=2e8, reps=1, log=lambda s: None)
    assert [c[0] for c in chosen] == [384, 1024]
    lines = [l.split("#")[0].split() for l in path.read_text().splitlines() if not l.startswith("#")]
    assert [list(map(int, l)) for l in lines] == [[w, c, v, r] for w, c, v, r, _ in chosen]

    jo, ro, b = synth.jobs(seed=20262, njobs=48)
    p = Params(adaptive=True)
    base = mh.batch_run(ctx, p, jo, ro, b)
    want = [(base.output(j, 0), base.output(j, 1)) for j in range(48)]
    old = os.environ.get("MPOA_TUNING")
    # A table that sends every width to a shape the built-in rule never
    # picks for these rows, so following it is visible in the results'
    # equality only through a different kernel.
    forced = tmp_path / "forced.txt"
    forced.write_text("100000 4 2 24\n")
    try:
        for t in (path, forced):
            os.environ["MPOA_TUNING"] = str(t)
            # The host-driven engine's launches and the device-resident
            # buckets (one length class each) both follow the table.
            for resident in (False, True):
                res = mh.batch_run(ctx, p, jo, ro, b, resident=resident)
                assert [(res.output(j, 0), res.output(j, 1)) for j in range(48)] == want
                assert res.stats()["cells"] == base.stats()["cells"]
    finally:
        if old is None:
            os.environ.pop("MPOA_TUNING", None)
        else:
            os.environ["MPOA_TUNING"] = old
