import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "minipoa-rs_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="module", params=["values", "records"])
def ctx(request):
    """A device context per DP layout: every GPU test module that takes
    `ctx` runs once with M/D/I values and once with traceback records in
    device memory (mpoa_ctx_set_layout); results must not differ."""
    import minipoa_hip as mh

    c = mh.Context(0)
    c.set_layout({"values": mh.LAYOUT_VALUES, "records": mh.LAYOUT_RECORDS}[request.param])
    yield c
    c.close()
