"""The committed golden fixtures (tests/golden/oracle_small.npz): the oracle still reproduces them,
and the product's host-side plane wave agrees with them. CPU only."""
import os
import sys

import numpy as np
import pytest

from conftest import rel

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden as MG  # noqa: E402


@pytest.fixture(scope="module")
def golden():
    return MG.load()


def test_oracle_reproduces_fixtures(oracle, golden):
    now = MG.compute()
    assert set(now) == set(golden)
    for k, v in golden.items():
        if v.dtype.kind in "iu":
            assert int(now[k]) == int(v), k
        else:
            assert rel(now[k], v) < 1e-12, (k, rel(now[k], v))


def test_fixture_solution_solves_the_system(oracle, golden):
    """The stored chi satisfies A chi = chi0 to the solver's tolerance (checked with the oracle)."""
    from extreme_amd import cartesian_model as cm
    m = MG.golden_model()
    w = cm.omega_of(m.frequencies[0])
    G = {c: golden[f"green_{c}"] for c in oracle.COMPONENTS}
    op = oracle.Operator(m, G)
    op.prepare("A", m.zeta(w).reshape(-1), m.zeta0(w))
    res = np.linalg.norm(op.apply(golden["chi"]) - golden["chi0"]) / np.linalg.norm(golden["chi0"])
    assert res < 1e-8
    assert golden["fgmres_bea"][-1] < 1e-8 <= golden["fgmres_bea"][-2]


def test_product_plane_wave_matches_fixture(golden):
    from extreme_amd import cartesian_model as cm
    from extreme_amd import forward as F
    sec = MG.golden_model().section
    for a, f in enumerate(golden["pw_freqs"]):
        w = cm.omega_of(float(f))
        for b, d in enumerate(golden["pw_depths"]):
            assert abs(F.plane_wave_e(sec, w, float(d)) / golden["pw_e"][a, b] - 1) < 1e-13
            assert abs(F.plane_wave_h(sec, w, float(d)) / golden["pw_h"][a, b] - 1) < 1e-13


def test_product_scattered_current_matches_fixture(golden):
    from extreme_amd import cartesian_model as cm
    from extreme_amd import forward as F
    m = MG.golden_model()
    w = cm.omega_of(m.frequencies[0])
    js = F.scattered_current(m, w, 0, F.normal_field(m, w))
    assert rel(js, golden["js"]) < 1e-13
