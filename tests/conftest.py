import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; calls the HIP path through the C ABI")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def graphs_golden():
    z = np.load(os.path.join(GOLDEN, "graphs_golden.npz"))
    sets = []
    for s in range(int(z["count"])):
        sets.append({
            "syms": z[f"s{s}_syms"],
            "min_ed": int(z[f"s{s}_min_ed"]),
            "T": {T: (z[f"s{s}_T{T}_row_ptr"], z[f"s{s}_T{T}_col"], z[f"s{s}_T{T}_codebook"]) for T in (1, 4)},
        })
    return sets


@pytest.fixture(scope="session")
def gpu_ctx():
    import indexgen_amd as igc
    print(f"[gpu] backend={igc.backend()}", flush=True)
    ctx = igc.Context(0)
    yield ctx
    ctx.close()
