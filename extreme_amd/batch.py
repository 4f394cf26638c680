"""Batch scheduler for many conductivity models: group by (grid, background, frequency), share the
Green tensor within a group, and keep several groups in flight on the GPU.

The reference solves one model and frequency after another (Mt3DForwardSolver.cs:78-129) and reuses
a Green tensor between models only through its TensorCache (Extreme.Cartesian/Forward/
TensorCache.cs:8-35). Here every group of up to `group_size` models that share the grid, the
background and the frequency becomes one ModelGroupSolve: one Green tensor, one Chi0 apply and one
batched FGMRES whose every operator apply serves all 2 x models right-hand sides (both
polarisations), each following exactly the iteration of its own solve. Groups are independent, so
`workers` host threads each drive their own operator handles; their kernels share the GPU, which
small grids alone leave mostly idle.
"""
from __future__ import annotations

import threading
import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from . import cartesian_model as cm
from . import forward as F
from . import xtreme as X


def _key(m: cm.CartesianModel, f: float):
    return (m.nx, m.ny, m.nz, float(m.dx), float(m.dy), tuple(np.round(m.depth, 6)), tuple(np.round(m.thick, 6)),
            tuple(m.section.thickness), tuple(m.section.sigma), float(m.section.zero_level), float(f))


def make_groups(models, group_size: int = 4):
    """(frequency, [model indices]) groups: each model at each of its frequencies, models sharing the
    grid, background and frequency together, at most group_size per group, in first-seen order."""
    byk = OrderedDict()
    for i, m in enumerate(models):
        for f in m.frequencies:
            byk.setdefault(_key(m, f), []).append(i)
    groups = []
    for k, idx in byk.items():
        for s in range(0, len(idx), group_size):
            groups.append((k[-1], idx[s:s + group_size]))
    return groups


@dataclass
class BatchResult:
    solves: int = 0                          # right-hand sides solved (models x frequencies x 2)
    groups: int = 0
    seconds: float = 0.0                     # wall time of the whole batch
    green_seconds: float = 0.0               # summed Green-tensor assembly time over the groups
    solve_seconds: float = 0.0               # summed FGMRES time over the groups
    converged: bool = True
    max_true_residual: float = 0.0
    iterations: dict = field(default_factory=dict)   # (model index, frequency) -> [pol X, pol Y]


def run(models, settings: X.ForwardSettings | None = None, group_size: int = 4, workers: int = 4,
        log=None, precision: str = "fp64") -> BatchResult:
    """Solves both polarisations of every model at each of its frequencies; precision="mixed"
    solves each group by FP64 refinement around complex64 corrections (ModelGroupSolve.solve_mixed),
    to the same FP64 true residual."""
    settings = settings or X.ForwardSettings()
    groups = make_groups(models, group_size)
    res = BatchResult(groups=len(groups))
    lock = threading.Lock()
    nxt = [0]
    errors = []

    def worker():
        import torch
        while True:
            with lock:
                if nxt[0] >= len(groups) or errors:
                    return
                g = nxt[0]
                nxt[0] += 1
            f, idx = groups[g]
            try:
                t0 = time.perf_counter()
                grp = F.ModelGroupSolve([models[i] for i in idx], f, settings, precision=precision)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                infos = grp.solve_mixed() if precision == "mixed" else grp.solve()
                t2 = time.perf_counter()
                with lock:
                    res.green_seconds += t1 - t0
                    res.solve_seconds += t2 - t1
                    res.solves += len(infos)
                    for r, i in enumerate(idx):
                        a, b = infos[2 * r], infos[2 * r + 1]
                        if precision == "mixed":
                            res.iterations[(i, f)] = [a.inner_iterations, b.inner_iterations]
                            res.converged = res.converged and a.converged and b.converged
                            res.max_true_residual = max(res.max_true_residual, a.residual, b.residual)
                        else:
                            res.iterations[(i, f)] = [a.number_of_iterations, b.number_of_iterations]
                            res.converged = res.converged and a.convergence_achieved and b.convergence_achieved
                            res.max_true_residual = max(res.max_true_residual, a.backward_error, b.backward_error)
                del grp
                if log:
                    log(f"group {g}: {len(idx)} models at {f:g} Hz ({models[idx[0]].nx}x{models[idx[0]].ny}x"
                        f"{models[idx[0]].nz}) in {t2 - t0:.3f} s")
            except Exception as e:   # stop the other workers, re-raised below
                with lock:
                    errors.append(e)
                return

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker) for _ in range(max(1, workers))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    res.seconds = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return res
