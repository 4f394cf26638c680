"""Training and evaluation of the learned right preconditioner (no counterpart in the reference).

The flexible FGMRES of FlexibleGmresWithGmresPreconditioner (Extreme.Fgmres/
FlexibleGmresWithGmresPreconditioner.cs:29-58) calls its right preconditioner once per outer
iteration (FlexibleGmresSolver.ApplyPreconditionerRight, FlexibleGmresSolver.cs:24-29): with the
reference's inner GMRES, z_j ~ A^-1 v_j from ten operator applies. Those pairs (v_j, z_j), harvested
from the solver's own iterations on seeded models, are the training data. The network
(LearnedPreconditioner: a 3x3x3 convolution pair plus a spectral branch mixing the 3 Nz complex
channels of each low lateral mode of the zero-padded spectrum) is trained on the GPU by the library's
own forward, backward and Adam kernels, then stands in for the inner GMRES; the FP64 true residual
still judges convergence.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import cartesian_model as cm
from . import forward as F
from . import xtreme as X


@dataclass
class PairSet:
    """The preconditioner pairs of one solve: model, frequency, polarisation, and the device vectors."""
    model: cm.CartesianModel
    frequency: float
    pol: int
    vs: list
    zs: list


def harvest(models, frequencies, settings: X.ForwardSettings, capacity: int = 16, pols=(0, 1)) -> list:
    """Solves every (model, frequency, polarisation) with FGMRES / inner GMRES on the GPU and keeps the
    first `capacity` preconditioner pairs of each solve."""
    out = []
    for m in models:
        for f in frequencies:
            p = F.FrequencySolve(m, f, settings)
            solver = X.AnomalyCurrentFgmresSolver(p.op, settings)
            for pol in pols:
                b = p.chi0[pol].contiguous()
                _, vs, zs = solver.solve_harvest(b, b.clone(), capacity)
                out.append(PairSet(m, f, pol, vs, zs))
            del p, solver
    return out


def _features(net, ps: PairSet):
    net.set_features(ps.model.sigma, ps.frequency)


def mean_loss(net, sets) -> float:
    """Mean relative error of the network's correction over all pairs."""
    tot, cnt = 0.0, 0
    for ps in sets:
        _features(net, ps)
        tot += net.gradients(ps.vs, ps.zs)[2] * len(ps.vs)
        cnt += len(ps.vs)
    return tot / max(cnt, 1)


def train(net, sets, steps: int = 400, lr: float = 1e-3, batch: int = 32, seed: int = 1, log=None) -> list:
    """Adam steps on batches of up to `batch` pairs, each batch from one solve (the network's
    features are that model's and frequency's). Returns the loss before each step."""
    rng = np.random.default_rng(seed)
    losses = []
    for step in range(steps):
        ps = sets[int(rng.integers(len(sets)))]
        _features(net, ps)
        idx = rng.choice(len(ps.vs), size=min(batch, len(ps.vs)), replace=False)
        losses.append(net.train_step([ps.vs[i] for i in idx], [ps.zs[i] for i in idx], lr=lr))
        if log and (step % 100 == 0 or step == steps - 1):
            log(f"step {step}: loss {losses[-1]:.4f}")
    return losses


@dataclass
class Trial:
    name: str
    converged: bool
    outer_iterations: int
    applies: int
    true_residual: float
    seconds: float


def compare(model, frequency: float, settings: X.ForwardSettings, nets: dict, pol: int = 0, repeats: int = 1) -> list:
    """One polarisation of one model and frequency solved with the reference's inner GMRES, with each
    network in `nets` (name -> LearnedPreconditioner) and with no preconditioner (the untrained
    network, which is the identity). Each run is timed after a warm-up run; all start from chi0."""
    import torch
    p = F.FrequencySolve(model, frequency, settings)
    solver = X.AnomalyCurrentFgmresSolver(p.op, settings)
    b = p.chi0[pol].contiguous()
    ident = X.LearnedPreconditioner(model.nx, model.ny, model.nz)
    runs = [("inner GMRES(10)", lambda x: solver.solve(b, x))]
    for name, net in nets.items():
        def run(x, net=net):
            net.set_features(model.sigma, frequency)
            return solver.solve_learned(b, x, net)
        runs.append((name, run))
    runs.append(("none (identity)", lambda x: solver.solve_learned(b, x, ident)))
    out = []
    for name, run in runs:
        run(b.clone())                              # warm-up: tile measurement, workspaces
        best = None
        for _ in range(repeats):
            x = b.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = run(x)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        out.append(Trial(name, info.convergence_achieved, info.number_of_iterations, info.matvecs,
                         info.backward_error, best))
    del p, solver
    return out
