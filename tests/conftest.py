import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

CONFIGS = ROOT / "tissuemc_amd" / "configs"
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def gpu():
    """The engine module, on a machine with a GPU. The library must load: a missing build fails."""
    import tissuemc_amd.engine as engine

    engine.load()
    if not engine.available():
        pytest.skip("no HIP device visible")
    return engine


def load_spec(name: str) -> dict:
    import oracle

    return oracle.parse_mci_text((CONFIGS / f"{name}.mci").read_text())[0]
