"""GPU FGMRES (K1-K13) against the oracle's restatement of Extreme.Fgmres on the same operator."""
import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import xtreme as X
from tests.conftest import rel

pytestmark = pytest.mark.gpu

TOL_SOLUTION = 1e-9   # FP64 GPU path within 1e-9 relative of the FP64 oracle


def _rhs(oracle, m, w, G, pol):
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    oop = oracle.Operator(m, G)
    e0 = oracle.normal_field(m, w)
    js = oracle.jscattered(m, pol, z, z0, e0)
    oop.prepare("Chi0", z, z0)
    chi0 = oop.apply(js)
    oop.prepare("A", z, z0)
    return oop, chi0


@pytest.mark.parametrize("m_outer,m_inner,max_repeat", [(30, 10, 100), (4, 3, 50), (2, 2, 3)])
def test_fgmres_matches_oracle_c1(oracle, torch_cuda, m_outer, m_inner, max_repeat):
    m = cm.config_c1()
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w)
    oop, chi0 = _rhs(oracle, m, w, G, 0)
    x_ref, info_ref, hist_ref = oracle.fgmres(oop.raw_apply, chi0, chi0, m_outer, m_inner, max_repeat, 1e-8)
    s = X.ForwardSettings(residual=1e-8, inner_buffer_length=m_inner, outer_buffer_length=m_outer,
                          max_repeats_number=max_repeat)
    op = X.ConvolutionOperator(m, w, s)
    op.set_green_tensor(G)
    op.prepare_operator("A", m.zeta(w))
    solver = X.AnomalyCurrentFgmresSolver(op, s)
    x = chi0.copy()
    info = solver.solve(chi0, x)
    print(f"m={m_outer}/{m_inner}/{max_repeat}: gpu it {info.number_of_iterations} mv {info.matvecs} "
          f"bea {info.arnoldi_backward_error:.3e} be {info.backward_error:.3e}; oracle it {info_ref['iterations']} "
          f"mv {info_ref['matvecs']}; x rel {rel(x, x_ref):.3e}", flush=True)
    assert info.number_of_iterations == info_ref["iterations"]
    assert info.matvecs == info_ref["matvecs"]
    assert info.convergence_achieved == info_ref["converged"]
    beas = [h[2] for h in solver.history]
    beas_ref = [h[2] for h in hist_ref if h[0] == 1]
    assert len(beas) == len(beas_ref)
    for a, b in zip(beas, beas_ref):
        assert abs(a - b) <= 1e-9 * max(b, 1e-12) + 1e-13
    assert rel(x, x_ref) < TOL_SOLUTION


def test_batch_solve_matches_single(oracle, torch_cuda):
    torch = torch_cuda
    m = cm.config_c1()
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w)
    s = X.ForwardSettings()
    op = X.ConvolutionOperator(m, w, s, max_rhs=2)
    op.set_green_tensor(G)
    rhs = [_rhs(oracle, m, w, G, pol)[1] for pol in (0, 1)]
    op.prepare_operator("A", m.zeta(w))
    solver = X.AnomalyCurrentFgmresSolver(op, s)
    singles = []
    for b in rhs:
        x = b.copy()
        info = solver.solve(b, x)
        singles.append((x, info))
    bs = [torch.from_numpy(b).to("cuda") for b in rhs]
    xs = [b.clone() for b in bs]
    infos = solver.solve_batch(bs, xs)
    for (x1, i1), xb, ib in zip(singles, xs, infos):
        print(f"batch: it {ib.number_of_iterations} vs {i1.number_of_iterations}, rel {rel(xb.cpu().numpy(), x1):.3e}",
              flush=True)
        assert ib.number_of_iterations == i1.number_of_iterations
        assert ib.matvecs == i1.matvecs
        assert rel(xb.cpu().numpy(), x1) < 1e-12


def test_zero_rhs_returns_immediately(oracle, torch_cuda):
    """bn == 0 -> ResultInfo.Converged(0, 0, 0) without touching x (Gmres.cs:80-83)."""
    m = cm.small_model(4, 4, 2)
    w = cm.omega_of(1.0)
    G = oracle.green_atoa(m, w)
    op = X.ConvolutionOperator(m, w)
    op.set_green_tensor(G)
    op.prepare_operator("A", m.zeta(w))
    solver = X.AnomalyCurrentFgmresSolver(op)
    b = np.zeros(m.n, dtype=np.complex128)
    x = np.ones(m.n, dtype=np.complex128)
    info = solver.solve(b, x)
    assert info.convergence_achieved and info.number_of_iterations == 0 and info.matvecs == 0
    assert np.all(x == 1)


def test_non_convergence_matches_oracle(oracle, torch_cuda):
    """FGMRES(2) / GMRES(1) with one repeat cannot reach 1e-12: ResultInfo.NonConverged with the
    oracle's iterations, matvecs, Arnoldi and true backward errors (Gmres.cs:131-141), and the same
    iterate."""
    m = cm.config_c1()
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w)
    oop, chi0 = _rhs(oracle, m, w, G, 1)
    x_ref, info_ref, _ = oracle.fgmres(oop.raw_apply, chi0, chi0, 2, 1, 1, 1e-12)
    assert not info_ref["converged"]
    s = X.ForwardSettings(residual=1e-12, inner_buffer_length=1, outer_buffer_length=2, max_repeats_number=1)
    op = X.ConvolutionOperator(m, w, s)
    op.set_green_tensor(G)
    op.prepare_operator("A", m.zeta(w))
    x = chi0.copy()
    info = X.AnomalyCurrentFgmresSolver(op, s).solve(chi0, x)
    print(f"non-converged: it {info.number_of_iterations} mv {info.matvecs} bea {info.arnoldi_backward_error:.4e} "
          f"be {info.backward_error:.4e}; oracle {info_ref}", flush=True)
    assert not info.convergence_achieved
    assert (info.number_of_iterations, info.matvecs) == (info_ref["iterations"], info_ref["matvecs"])
    assert abs(info.arnoldi_backward_error - info_ref["bea"]) <= 1e-9 * info_ref["bea"]
    assert abs(info.backward_error - info_ref["be"]) <= 1e-9 * info_ref["be"]
    assert rel(x, x_ref) < TOL_SOLUTION


def test_exact_initial_guess_returns_at_once(oracle, torch_cuda):
    """b = A x0 with the same device apply: the first residual is exactly zero, so the solve returns
    x0 with ResultInfo.Converged(0, 0, 1) after one matvec (Gmres.cs:95-107)."""
    m = cm.small_model(4, 4, 2, seed=9)
    w = cm.omega_of(1.0)
    G = oracle.green_atoa(m, w)
    op = X.ConvolutionOperator(m, w)
    op.set_green_tensor(G)
    op.prepare_operator("A", m.zeta(w))
    rng = np.random.default_rng(2)
    x0 = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    b = op.apply(x0)
    x = x0.copy()
    info = X.AnomalyCurrentFgmresSolver(op).solve(b, x)
    print(f"exact guess: converged {info.convergence_achieved} it {info.number_of_iterations} mv {info.matvecs}",
          flush=True)
    assert info.convergence_achieved and info.number_of_iterations == 1 and info.matvecs == 1
    assert info.arnoldi_backward_error == 0 and info.backward_error == 0
    assert np.array_equal(x, x0)


def test_backward_error_exception(oracle, torch_cuda):
    """BackwardErrorChecking.CheckWithException: a solve that stops above the tolerance raises
    BackwardErrorException with "be = ..., tolerance = ..." (Gmres.cs:217-223), here an error
    status whose message carries the same text."""
    m = cm.small_model(4, 4, 2, seed=9)
    w = cm.omega_of(1.0)
    G = oracle.green_atoa(m, w)
    s = X.ForwardSettings(residual=1e-14, inner_buffer_length=1, outer_buffer_length=1, max_repeats_number=1)
    op = X.ConvolutionOperator(m, w, s)
    op.set_green_tensor(G)
    op.prepare_operator("A", m.zeta(w))
    rng = np.random.default_rng(4)
    b = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    x = b.copy()
    with pytest.raises(RuntimeError, match=r"be = .*, tolerance = 1e-14"):
        X.AnomalyCurrentFgmresSolver(op, s, backward_error_checking=2).solve(b, x)
    x = b.copy()
    info = X.AnomalyCurrentFgmresSolver(op, s, backward_error_checking=1).solve(b, x)   # warning only
    assert not info.convergence_achieved and info.backward_error >= 1e-14


def test_non_finite_operator_is_refused(oracle, torch_cuda):
    """A Green tensor holding NaN (what the layer recursions give when exp(eta h) overflows) makes
    every Krylov vector NaN; the solve must raise rather than stop as if converged (NaN > tol is
    false in the stopping test)."""
    m = cm.small_model(4, 4, 2, seed=9)
    w = cm.omega_of(1.0)
    G = oracle.green_atoa(m, w)
    G["xx"] = G["xx"].copy()
    G["xx"][3, 0] = np.nan
    op = X.ConvolutionOperator(m, w)
    op.set_green_tensor(G)
    op.prepare_operator("A", m.zeta(w))
    rng = np.random.default_rng(4)
    b = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    x = b.copy()
    with pytest.raises(RuntimeError, match="NaN or Inf"):
        X.AnomalyCurrentFgmresSolver(op).solve(b, x)
