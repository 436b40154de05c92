"""Oracle against the reference compiled from its sources (oracle/_ref), where that build exists."""
import os

import numpy as np
import pytest

from conftest import normwise
from flint_amd import states as S

REF_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libflint_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference build (make -C oracle ref) absent")


@pytest.mark.parametrize("mech,seed", [("WD", 101), ("TSR-GP-24", 102), ("TSR-Rich-31", 103)])
def test_oracle_equals_reference_build(oracle_mod, mechs, mech, seed):
    md = mechs[mech]
    if mech == "WD":
        st = S.wd_states(md.species, md.W, seed, 0, 3000)
    else:
        st = S.radical_states(md.species, md.W, seed, 0, 3000, T_lo=600.0, T_hi=4000.0, ignition_half=True)
    rho = S.soa(st["rho"])
    w, sc = oracle_mod.Oracle(md).wdot(rho, st["T"])
    wr = oracle_mod.Reference(md).wdot(rho, st["T"])
    assert normwise(w, wr, sc) <= 1e-14
