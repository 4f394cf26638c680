"""Mixed precision (complex64 operator inside FP64 iterative refinement) against the FP64 oracle:
the FP64 true residual reaches the tolerance and the solution's error is stated per result."""
import numpy as np
import pytest
import torch

from extreme_amd import cartesian_model as cm
from extreme_amd import forward as F
from extreme_amd import xtreme as X
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def test_mixed_c1_reaches_fp64_true_residual(oracle, torch_cuda):
    m = cm.config_c1()
    p = F.FrequencySolve(m, m.frequencies[0], precision="mixed")
    infos = p.solve_mixed()
    chi_mixed = p.chi.clone()
    w = p.omega
    G = p.op.green_tensor()
    oop = oracle.Operator(m, G)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    for pol in (0, 1):
        chi0 = p.chi0[pol].cpu().numpy()
        oop.prepare("A", z, z0)
        x = chi_mixed[pol].cpu().numpy()
        true_res = np.linalg.norm(chi0 - oop.apply(x)) / np.linalg.norm(chi0)
        x_ref, info, _ = oracle.fgmres(oop.raw_apply, chi0, chi0, 30, 10, 100, 1e-8)
        err = rel(x, x_ref)
        print(f"pol {pol}: refinements {infos[pol].refinements} fp32 applies {infos[pol].matvecs32} "
              f"fp64 applies {infos[pol].matvecs64} residual {infos[pol].residual:.2e} "
              f"(oracle-checked {true_res:.2e}); |x - x_fp64| / |x_fp64| = {err:.2e}", flush=True)
        assert infos[pol].converged
        assert infos[pol].residual <= 1e-8
        assert true_res <= 1.05e-8
        assert abs(true_res - infos[pol].residual) <= 1e-3 * true_res + 1e-12
        assert err < 1e-6


def test_mixed_operator_fp64_applies_unchanged(oracle, torch_cuda):
    """An XTR_MIXED handle applies in FP64 outside the refined solve: same result as XTR_FP64."""
    m = cm.small_model(8, 6, 4, seed=9)
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w)
    ops = {}
    for prec in ("fp64", "mixed"):
        ops[prec] = X.ConvolutionOperator(m, w, precision=prec)
        ops[prec].set_green_tensor(G)
        ops[prec].prepare_operator("A")
    x = np.random.default_rng(4).standard_normal(m.n) + 0.25j
    assert np.array_equal(ops["fp64"].apply(x), ops["mixed"].apply(x))
    # the tensor is stored folded; the export rebuilds the mirrored blocks from their sources, which
    # the oracle's FFT (length 12 here) produces equal up to rounding
    assert ops["mixed"].green_layout()[0] == "folded"
    back = ops["mixed"].green_tensor()
    for c in X.ConvolutionOperator.COMPONENTS:
        assert np.abs(back[c] - G[c]).max() <= 1e-14 * np.abs(G[c]).max()
