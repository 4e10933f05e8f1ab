import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


def golden_index():
    with open(os.path.join(GOLDEN, "index.json")) as f:
        return json.load(f)


def load_golden(case: str) -> dict:
    with np.load(os.path.join(GOLDEN, case + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_pairs(gd: dict) -> np.ndarray:
    """The input edges of a golden case: stored, or regenerated from the
    seeded generator the case names (checked against its SHA-256)."""
    if "pairs" in gd:
        return gd["pairs"]
    import hashlib
    from kclistpp_amd import graphgen as gg
    part, seed = (int(x) for x in gd["gen"])
    pairs = gg.bipartite_plus(part, float(gd["p_in"]), seed)
    assert hashlib.sha256(pairs.tobytes()).hexdigest() == str(gd["pairs_sha"][0])
    return pairs


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def kcl():
    import kclistpp_amd
    kclistpp_amd.lib()
    return kclistpp_amd
