"""Several models sharing one Green tensor (batched right-hand sides with their own R): applies and
solves equal the one-model-at-a-time results, and the streaming contraction with 3 and 4
right-hand sides equals the generic kernel."""
import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import forward as F
from extreme_amd import xtreme as X
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def _group(nx, ny, nz, k, seed0=10):
    base = cm.small_model(nx, ny, nz, seed=seed0)
    out = [base]
    for i in range(1, k):
        m = cm.small_model(nx, ny, nz, seed=seed0 + i)     # same grid and background, other sigma
        out.append(m)
    return out


@pytest.mark.parametrize("shape,k", [((6, 4, 3), 2), ((6, 8, 12), 4), ((5, 4, 24), 4), ((12, 6, 32), 2), ((12, 6, 32), 3)])
def test_models_apply_equals_single_model(oracle, torch_cuda, shape, k):
    torch = torch_cuda
    models = _group(*shape, k)
    w = cm.omega_of(models[0].frequencies[0])
    G = oracle.green_atoa(models[0], w)
    op = X.ConvolutionOperator(models[0], w, max_rhs=2 * k)
    op.set_green_tensor(G)
    rng = np.random.default_rng(3)
    xs = [torch.from_numpy(rng.standard_normal(models[0].n) + 1j * rng.standard_normal(models[0].n)).to("cuda")
          for _ in range(2 * k)]
    rhs_model = [i // 2 for i in range(2 * k)]
    for kind in ("A", "Chi0"):
        op.prepare_models(kind, [m.zeta(w) for m in models])
        ys = op.apply_models(xs, rhs_model)
        for r in range(2 * k):
            oop = oracle.Operator(models[rhs_model[r]], G)
            oop.prepare(kind, models[rhs_model[r]].zeta(w).reshape(-1), models[rhs_model[r]].zeta0(w))
            e = rel(ys[r].cpu().numpy(), oop.apply(xs[r].cpu().numpy()))
            assert e < 1e-12, (kind, r, e)


@pytest.mark.parametrize("nrhs", [3, 4])
def test_streaming_contraction_three_and_four_rhs(oracle, torch_cuda, monkeypatch, nrhs):
    torch = torch_cuda
    m = cm.small_model(24, 12, 32, seed=6)
    w = cm.omega_of(m.frequencies[0])
    rng = np.random.default_rng(nrhs)
    L, S = 4 * m.nx * m.ny, 32 * 33 // 2
    G = {c: rng.standard_normal((L, 1024 if c in ("xz", "yz") else S))
         + 1j * rng.standard_normal((L, 1024 if c in ("xz", "yz") else S)) for c in oracle.COMPONENTS}
    op = X.ConvolutionOperator(m, w, max_rhs=4)
    op.set_green_tensor(G)
    op.prepare_operator("A")
    xs = rng.standard_normal((nrhs, m.n)) + 1j * rng.standard_normal((nrhs, m.n))
    y = op.apply(xs)
    monkeypatch.setenv("XTR_CONTRACT_KERNEL", "generic")
    yg = op.apply(xs)
    monkeypatch.delenv("XTR_CONTRACT_KERNEL")
    oop = oracle.Operator(m, G)
    oop.prepare("A", m.zeta(w).reshape(-1), m.zeta0(w))
    print(f"nrhs {nrhs}: stream vs generic {rel(y, yg):.2e}, vs oracle {rel(y[-1], oop.apply(xs[-1])):.2e}", flush=True)
    assert rel(y, yg) < 1e-14
    assert rel(y[-1], oop.apply(xs[-1])) < 1e-12


def test_model_group_solve_equals_separate_solves(oracle, torch_cuda):
    """Two models x two polarisations in one batched solve: each right-hand side takes the same
    iterations and reaches the same solution as its model's own solve."""
    models = _group(8, 8, 4, 2, seed0=30)
    f = models[0].frequencies[0]
    grp = F.ModelGroupSolve(models, f)
    infos = grp.solve()
    for k, m in enumerate(models):
        p = F.FrequencySolve(m, f)
        single = p.solve()
        for pol in (0, 1):
            r = 2 * k + pol
            e = rel(grp.chi[r].cpu().numpy(), p.chi[pol].cpu().numpy())
            print(f"model {k} pol {pol}: its {infos[r].number_of_iterations} vs {single[pol].number_of_iterations}, "
                  f"rel {e:.2e}", flush=True)
            assert infos[r].number_of_iterations == single[pol].number_of_iterations
            assert infos[r].matvecs == single[pol].matvecs
            assert e < 1e-12


def test_model_group_mixed_solve(oracle, torch_cuda):
    """Two models x two polarisations in one mixed-precision batched solve (FP64 refinement around
    complex64 corrections, each right-hand side with its own model's R): every FP64 true residual
    reaches 1e-8, checked again with the oracle's operator of that model, and each solution is within
    1e-6 of the FP64 group solve."""
    models = _group(8, 8, 4, 2, seed0=30)
    f = models[0].frequencies[0]
    w = cm.omega_of(f)
    grp = F.ModelGroupSolve(models, f, precision="mixed")
    infos = grp.solve_mixed()
    ref = F.ModelGroupSolve(models, f)
    ref.solve()
    G = oracle.green_atoa(models[0], w)
    for r in range(4):
        m = models[grp.rhs_model[r]]
        oop = oracle.Operator(m, G)
        oop.prepare("A", m.zeta(w).reshape(-1), m.zeta0(w))
        b = grp.chi0[r].cpu().numpy()
        x = grp.chi[r].cpu().numpy()
        res = rel(oop.apply(x), b)
        e = rel(x, ref.chi[r].cpu().numpy())
        print(f"rhs {r}: refinements {infos[r].refinements}, complex64 applies {infos[r].matvecs32}, "
              f"residual {infos[r].residual:.2e} (oracle {res:.2e}), vs FP64 {e:.2e}", flush=True)
        assert infos[r].converged and infos[r].residual <= 1e-8
        assert res <= 2e-8
        assert e < 1e-6
