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


def _oracle_responses(oracle, m, w, chis, level_z=0.0):
    """E, H, Z, rho_a and phase at the level from the oracle's AtoO path (AtoOCalculator.cs:63-160,
    Mt3DForwardSolver.cs:236-271, ResponseFunctionsCalculator.cs:36-51) for the two polarisations'
    solutions chis."""
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    Ge = oracle.green_atoo_level(m, w, level_z, True)
    Gh = oracle.green_atoo_level(m, w, level_z, False)
    e0 = oracle.plane_wave_e(m.section, w, level_z)
    h0 = oracle.plane_wave_h(m.section, w, level_z)
    en = oracle.normal_field(m, w)
    E, H = [], []
    for pol in (0, 1):
        js = oracle.jscattered(m, pol, z, z0, en)
        _, jq = oracle.escattered_jq(m, z, z0, chis[pol], js)
        ea = oracle.atoo_field(m, Ge, True, jq)
        ha = oracle.atoo_field(m, Gh, False, jq)
        E.append(ea + np.array([e0 if k == pol else 0 for k in range(3)]))
        H.append(ha + (np.array([0, h0, 0]) if pol == 0 else np.array([-h0, 0, 0])))
    Z = F.impedance(E[0], E[1], H[0], H[1])
    rho, phase = F.apparent_resistivity_phase(Z, w)
    return {"E": E, "H": H, "Z": Z, "rho": rho, "phase": phase}


@pytest.mark.parametrize("vectors", ["fp32", "fp64"])
def test_mixed_c1_reaches_fp64_true_residual(oracle, torch_cuda, vectors):
    m = cm.config_c1()
    p = F.FrequencySolve(m, m.frequencies[0], precision="mixed")
    infos = p.solve_mixed(vectors=vectors)
    chi_mixed = p.chi.clone()
    w = p.omega
    G = p.op.green_tensor()
    oop = oracle.Operator(m, G)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    refs = []
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
        refs.append(x_ref)
    # every observed result of the mixed solve against the oracle's FP64 solve, per quantity
    p.build_level(0.0)
    got = p.level_responses()
    errs = F.result_errors(got, _oracle_responses(oracle, m, w, refs))
    print(f"{vectors} vectors: result errors vs the FP64 oracle {errs}", flush=True)
    for q, e in errs.items():
        assert e < 1e-6, (q, e)


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


def test_complex64_vector_apply(oracle, torch_cuda):
    """The complex64 apply on complex64 vectors (the correction solves' operator) against the FP64
    apply of the same handle: within the complex64 rounding of the tensor, FFTs and vectors."""
    torch = torch_cuda
    m = cm.config_c1()
    p = F.FrequencySolve(m, m.frequencies[0], precision="mixed")
    x = p.chi0.clone()
    y64 = p.op.apply(x)
    y32 = p.op.apply_c64(x.to(torch.complex64))
    e = float(torch.linalg.norm(y32.to(torch.complex128) - y64) / torch.linalg.norm(y64))
    print(f"complex64 apply on complex64 vectors: rel error {e:.2e} against FP64", flush=True)
    assert e < 2e-6
    # the handle goes back to FP64 applies afterwards
    assert torch.equal(p.op.apply(x), y64)
