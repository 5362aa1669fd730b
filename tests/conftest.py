import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def engine():
    if not _has_gpu():
        pytest.skip("no GPU visible")
    from bespokeolap_artifacts_amd.engine import Engine
    eng = Engine(0)
    yield eng
    eng.close()


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # Names the case on stdout (flushed) before any GPU work, so a run that
    # dies inside a kernel still says which test it was in.
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
