"""The HIP library against the committed golden fixtures (tests/golden/oracle_small.npz): device
Green tensor, A and Chi0 applies, and the FGMRES solve with its iteration history."""
import os
import sys

import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import xtreme as X
from tests.conftest import rel

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden as MG  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-9   # FP64 GPU path within 1e-9 relative of the FP64 oracle


@pytest.fixture(scope="module")
def case(torch_cuda):
    g = MG.load()
    m = MG.golden_model()
    w = cm.omega_of(m.frequencies[0])
    s = X.ForwardSettings()
    op = X.ConvolutionOperator(m, w, s, max_rhs=2)
    op.calculate_green_tensor()
    return g, m, w, s, op


def test_device_green_matches_fixture(case):
    g, m, w, s, op = case
    got = op.green_tensor()
    for c in X.ConvolutionOperator.COMPONENTS:
        e = rel(got[c], g[f"green_{c}"])
        print(f"green {c}: rel {e:.2e}", flush=True)
        assert e < 1e-12, c


@pytest.mark.parametrize("kind", ["A", "Chi0"])
def test_apply_matches_fixture(case, kind):
    g, m, w, s, op = case
    op.prepare_operator(kind)
    y = op.apply(g["x"])
    e = rel(y, g["y_a" if kind == "A" else "y_chi0"])
    print(f"{kind}: rel {e:.2e}", flush=True)
    assert e < 1e-12


def test_fgmres_matches_fixture(case):
    g, m, w, s, op = case
    op.prepare_operator("Chi0")
    chi0 = op.apply(g["js"])
    assert rel(chi0, g["chi0"]) < 1e-12
    op.prepare_operator("A")
    solver = X.AnomalyCurrentFgmresSolver(op, s)
    x = chi0.copy()
    info = solver.solve(chi0, x)
    print(f"fgmres: it {info.number_of_iterations} mv {info.matvecs} x rel {rel(x, g['chi']):.2e}", flush=True)
    assert info.number_of_iterations == int(g["fgmres_iterations"])
    assert info.matvecs == int(g["fgmres_matvecs"])
    beas = np.array([h[2] for h in solver.history])
    assert beas.shape == g["fgmres_bea"].shape
    assert np.all(np.abs(beas - g["fgmres_bea"]) <= 1e-9 * g["fgmres_bea"] + 1e-13)
    assert rel(x, g["chi"]) < TOL


def test_level_tensors_and_fields_match_fixture(case, torch_cuda):
    torch = torch_cuda
    g, m, w, s, op = case
    for electric, tag in ((True, "e"), (False, "h")):
        op.calculate_green_atoo(0.0, electric)
        got = op.green_atoo(electric)
        scale = max(np.linalg.norm(g[f"atoo_{tag}_{c}"]) for c in got)
        for c, v in got.items():
            assert np.linalg.norm(v - g[f"atoo_{tag}_{c}"]) < 1e-11 * scale, (tag, c)
        field = op.level_field(torch.from_numpy(g["jq"]).to("cuda"), electric).cpu().numpy()
        print(f"level field {tag}: rel {rel(field, g[f'level_field_{tag}']):.2e}", flush=True)
        assert rel(field, g[f"level_field_{tag}"]) < 1e-10
