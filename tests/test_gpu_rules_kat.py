"""The hand-derived known answers of tests/rule_kats.py on the device: the
path of every case through the i32 and the i64 kernel (one attempt, no
retry), and the band geometry of the banded cases as the DP cells a served
job reports (the head/tail support ties and the adaptive recentring and
extra rules decide every row's band)."""
import pytest

import minipoa_hip as mh
import rule_kats as K

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", K.PATH_CASES, ids=lambda c: c.name)
@pytest.mark.parametrize("bits", [32, 64])
def test_device_path_known_answers(ctx, case, bits):
    g = case.build(mh.Graph)
    path = mh.poa_kernel(ctx, case.params, g, 0, 1, case.query, case.adaptive, case.band, bits)
    assert path == case.path


@pytest.mark.parametrize("case", [c for c in K.PATH_CASES if c.cells is not None],
                         ids=lambda c: c.name)
def test_device_band_geometry_known_answers(ctx, case):
    """One read aligned to the case's graph by a served job (align -> poa
    with the params' band): the job's DP cells are the rows' band widths."""
    g = case.build(mh.Graph)
    with mh.Server(ctx, case.params) as srv:
        t = srv.submit([case.query], graph=g, existing=0)
        srv.wait(t)
        assert srv.status(t) == 0, mh.last_error()
        assert srv.times(t)["cells"] == case.cells
