import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (HERE, os.path.join(REPO, "marmel-trace_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    g = os.path.join(HERE, "golden")
    return {name: np.load(os.path.join(g, name + ".npz"))
            for name in ("default_scene", "host_scenes", "vectors", "fog", "ingest", "adaptive", "flare",
                         "lens_texture")}
