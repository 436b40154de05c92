import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def mechs():
    from flint_amd.mechdata import mech_data
    return {n: mech_data(n) for n in ("WD", "TSR-GP-24", "TSR-Rich-31")}


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def golden_states(md, g):
    from flint_amd import states as S
    gen = str(g["gen"])
    if gen == "wd":
        return S.wd_states(md.species, md.W, int(g["seed"]), int(g["start"]), int(g["count"]))
    return S.radical_states(md.species, md.W, int(g["seed"]), int(g["start"]), int(g["count"]),
                            ignition_half=bool(g["ignition_half"]))


def normwise(x, ref, scale):
    """max_k,i |x - ref| / scale (scale = sum of |terms| of the reference sum)."""
    return float(np.max(np.abs(x - ref) / np.maximum(scale, np.finfo(float).tiny)))
