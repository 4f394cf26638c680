"""Host side of the MT forward solve around the GPU operator: the caller logic the reference's C# keeps.

Restates, in host FP64 (a few values per layer or per cell, not the hot path):
  PlaneWaveCalculator.CalculateFieldE/H  (Extreme.Cartesian.Magnetotellurics/PlaneWaveCalculator.cs:11-83)
  Mt3DForwardSolver.CalculateNormalFieldFromSource  (Solver/Mt3DForwardSolver.cs:251-275)
  ForwardSolver.CalculateJScattered / CalculateEScatteredFrom / CalculateJqFrom
                                          (Extreme.Cartesian/Forward/ForwardSolver.cs:186-449)
and drives the GPU for CalculateChi0From and SolveEquationFor (ForwardSolver.cs:253-307) with both
polarisations solved side by side (Mt3DForwardSolver.cs:95-129).
"""
from __future__ import annotations

import time

import numpy as np

from . import cartesian_model as cm
from . import xtreme as X


def _pw_container(section: cm.Section1D, omega: float, rc: int, tr: int):
    """eta, exp, P1, Q1 and A at lambda = 0 (PlanCalculator.cs:195-314)."""
    nl = section.nlayers
    zeta = section.zeta(omega)
    eta = np.array([np.sqrt(complex(0.0 - (1j * omega * zeta[k] * cm.MU0).real,
                                    0.0 - (1j * omega * zeta[k] * cm.MU0).imag)) for k in range(nl)])
    ex = np.exp(-2 * eta * np.asarray(section.thickness, dtype=float))
    q = np.zeros(nl, dtype=complex)
    n = nl - 1
    part = eta[n] / eta[n - 1]
    q[n - 1] = (1 - part) / (1 + part)
    for k in range(n - 1, 1, -1):
        qe = q[k] * ex[k]
        part = (eta[k] / eta[k - 1]) * ((qe - 1) / (qe + 1))
        q[k - 1] = (1 + part) / (1 - part)
    p = np.zeros(nl, dtype=complex)
    p[1] = (eta[1] - eta[0]) / (eta[1] + eta[0])
    for k in range(1, nl - 1):
        pe = p[k] * ex[k]
        part = (eta[k] / eta[k + 1]) * ((pe - 1) / (pe + 1))
        p[k + 1] = (1 + part) / (1 - part)
    a = np.zeros((nl, nl), dtype=complex)
    s, r = max(rc, tr), min(rc, tr)
    for nn in range(s, r - 1, -1):
        prev = 1.0 / (eta[nn] * (1 - p[nn] * q[nn] * ex[nn]))
        a[nn, nn] = prev
        for k in range(nn - 1, r - 1, -1):
            dk1 = section.layer_depth(k + 1)
            qq = ((1 + p[k + 1]) / (1 + p[k] * ex[k])) * np.exp((eta[k + 1] - eta[k]) * dk1)
            prev = qq * prev
            a[k, nn] = prev
            a[nn, k] = prev
    return eta, q, a


def plane_wave_e(section: cm.Section1D, omega: float, depth: float, thick: float = 0.0) -> complex:
    """PlaneWaveCalculator.CalculateFieldE (:11-40): normal E at depth (averaged over thick)."""
    tr = section.layer_at_depth(0.0)
    rc = section.layer_at_depth(depth)
    eta, q, a = _pw_container(section, omega, rc, tr)
    zeta = section.zeta(omega)
    ikr = 1j * np.sqrt(1j * omega * cm.MU0 * zeta[rc])
    ik1 = 1j * np.sqrt(1j * omega * cm.MU0 * zeta[tr])
    dr1 = section.layer_depth(rc + 1)
    d2 = section.layer_depth(tr + 1)
    aa = a[rc, tr] / a[tr, tr]
    if thick == 0:
        upper = np.exp(ikr * depth) + q[rc] * np.exp(ikr * (2 * dr1 - depth))
    else:
        z1, z2 = depth - thick / 2, depth + thick / 2
        u1 = np.exp(ikr * z1) - q[rc] * np.exp(ikr * (2 * dr1 - z1))
        u2 = np.exp(ikr * z2) - q[rc] * np.exp(ikr * (2 * dr1 - z2))
        upper = (u2 - u1) / (z2 - z1) / ikr
    lower = 1 + q[tr] * np.exp(ik1 * (2 * d2))
    return complex(aa * (upper / lower))


def plane_wave_h(section: cm.Section1D, omega: float, depth: float) -> complex:
    """PlaneWaveCalculator.CalculateFieldH (:42-64)."""
    tr = section.layer_at_depth(0.0)
    rc = section.layer_at_depth(depth)
    eta, q, a = _pw_container(section, omega, rc, tr)
    zeta = section.zeta(omega)
    kkr = 1j * omega * cm.MU0 * zeta[rc]
    ikr = 1j * np.sqrt(kkr)
    ik1 = 1j * np.sqrt(1j * omega * cm.MU0 * zeta[tr])
    dr1 = section.layer_depth(rc + 1)
    d2 = section.layer_depth(tr + 1)
    aa = a[rc, tr] / a[tr, tr]
    upper = np.exp(ikr * depth) - q[rc] * np.exp(ikr * (2 * dr1 - depth))
    lower = 1 + q[tr] * np.exp(ik1 * (2 * d2))
    return complex((np.sqrt(kkr) / (omega * cm.MU0)) * aa * (upper / lower))


def normal_field(model: cm.CartesianModel, omega: float) -> np.ndarray:
    """E0 per anomaly layer, averaged over the layer (Mt3DForwardSolver.cs:271-275)."""
    return np.array([plane_wave_e(model.section, omega, float(d + t / 2), float(t))
                     for d, t in zip(model.depth, model.thick)], dtype=np.complex128)


def scattered_current(model: cm.CartesianModel, omega: float, pol: int, e0: np.ndarray) -> np.ndarray:
    """j_s = (zeta - zeta0) E0 for polarisation pol (0 = X, 1 = Y), [Nx][Ny][3][Nz] (ForwardSolver.cs:225-237)."""
    nx, ny, nz = model.nx, model.ny, model.nz
    js = np.zeros((nx, ny, 3, nz), dtype=np.complex128)
    js[:, :, pol, :] = (model.zeta(omega) - model.zeta0(omega)[None, None, :]) * e0[None, None, :]
    return js.reshape(-1)


def scattered_field_and_jq(model: cm.CartesianModel, omega: float, chi: np.ndarray, js: np.ndarray):
    """E_s and jQ (ForwardSolver.cs:391-406, 424-436)."""
    nx, ny, nz = model.nx, model.ny, model.nz
    z = model.zeta(omega).reshape(nx, ny, 1, nz)
    z0 = model.zeta0(omega).reshape(1, 1, 1, nz)
    chi = np.asarray(chi).reshape(nx, ny, 3, nz)
    js = np.asarray(js).reshape(nx, ny, 3, nz)
    es = (2 * np.sqrt(z0.real + 0j) * chi - js) / (z + np.conj(z0))
    jq = es * (z - z0) + js
    return es.reshape(-1), jq.reshape(-1)


def impedance(e1, e2, h1, h2):
    """ResponseFunctionsCalculator.CalculateTensor (Magnetotellurics/ResponseFunctionsCalculator.cs:36-51):
    Z = [E1 E2] [H1 H2]^-1 per site from the two polarisations' field 3-vectors ([..., 3] -> [..., 2, 2])."""
    det = h1[..., 0] * h2[..., 1] - h2[..., 0] * h1[..., 1]
    hi11, hi12 = h2[..., 1] / det, -h2[..., 0] / det
    hi21, hi22 = -h1[..., 1] / det, h1[..., 0] / det
    z = np.empty(np.shape(e1)[:-1] + (2, 2), dtype=np.complex128)
    z[..., 0, 0] = e1[..., 0] * hi11 + e2[..., 0] * hi21
    z[..., 0, 1] = e1[..., 0] * hi12 + e2[..., 0] * hi22
    z[..., 1, 0] = e1[..., 1] * hi11 + e2[..., 1] * hi21
    z[..., 1, 1] = e1[..., 1] * hi12 + e2[..., 1] * hi22
    return z


def apparent_resistivity_phase(z, omega):
    """PlainTextExporter.CalculateApparentResistivity / ProcessPhase (Results/PlainTextExporter.cs:360-420):
    rho_a = |Z|^2 / (omega mu0) and phase = arg Z in radians (no degree conversion, no sign flip)."""
    return np.abs(z) ** 2 / (omega * cm.MU0), np.angle(z)


def result_errors(got: dict, ref: dict) -> dict:
    """Relative difference of each observed quantity between two level_responses() results, per
    quantity over all sites and both polarisations: |got - ref|_2 / |ref|_2 for E and H (full
    fields), and for Z, rho_a and phase over the off-diagonal elements xy and yx (the diagonal ones
    vanish over a 1-D earth, so their relative error is not meaningful)."""
    out = {}
    for q in ("E", "H"):
        a, b = np.stack(got[q]), np.stack(ref[q])
        out[q] = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    off = (Ellipsis, [0, 1], [1, 0])
    for q in ("Z", "rho", "phase"):
        a, b = got[q][off], ref[q][off]
        out[q] = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
    return out


def level_normal_fields(section: cm.Section1D, omega: float, level_z: float):
    """Mt3DForwardSolver.CalculateNormalFieldE/H(level) (Solver/Mt3DForwardSolver.cs:236-271): pol X
    has E = (E0, 0, 0), H = (0, H0, 0); pol Y has E = (0, E0, 0), H = (-H0, 0, 0)."""
    e0, h0 = plane_wave_e(section, omega, level_z), plane_wave_h(section, omega, level_z)
    return (np.array([e0, 0, 0]), np.array([0, h0, 0])), (np.array([0, e0, 0]), np.array([-h0, 0, 0]))


class FrequencySolve:
    """One frequency of Mt3DForwardSolver.SolvePrivate on the GPU: Green tensor, chi0 for both
    polarisations, and the side-by-side FGMRES solve of A chi = chi0."""

    def __init__(self, model: cm.CartesianModel, frequency: float, settings: X.ForwardSettings | None = None,
                 precision: str = "fp64"):
        import torch
        self.torch = torch
        self.model = model
        self.frequency = frequency
        self.omega = cm.omega_of(frequency)
        self.settings = settings or X.ForwardSettings()
        self.op = X.ConvolutionOperator(model, self.omega, self.settings, precision=precision, max_rhs=2)
        t0 = time.perf_counter()
        self.op.calculate_green_tensor()
        self.green_seconds = time.perf_counter() - t0
        self.atoo_green_seconds = 0.0
        self.atoo_field_seconds = [0.0, 0.0]
        t0 = time.perf_counter()
        e0 = normal_field(model, self.omega)
        self.js = [scattered_current(model, self.omega, pol, e0) for pol in (0, 1)]
        self.js_seconds = time.perf_counter() - t0               # CalcJScattered, both polarisations
        t0 = time.perf_counter()
        self.op.prepare_operator("Chi0")
        js_dev = torch.from_numpy(np.stack(self.js)).to("cuda")
        self.chi0 = self.op.apply(js_dev)                       # CalculateChi0From, both pols in one apply
        self.op.prepare_operator("A")
        torch.cuda.synchronize()
        self.chi0_seconds = time.perf_counter() - t0             # CalcChi0
        self.solver = X.AnomalyCurrentFgmresSolver(self.op, self.settings)
        self.chi = self.chi0.clone()
        self.level_z = None

    def solve(self):
        """SolveEquationFor for pol X and pol Y: chi <- chi0, then FGMRES (ForwardSolver.cs:300-307)."""
        self.chi.copy_(self.chi0)
        self.solver.history.clear()
        return self.solver.solve_batch([self.chi0[0], self.chi0[1]], [self.chi[0], self.chi[1]])

    def observe(self, level_z: float = 0.0, x_shift: float = 0.0, y_shift: float = 0.0) -> dict:
        """Fields and responses at the observation level after solve(): level tensors on the GPU,
        jQ from chi, anomalous E and H by the AtoO convolution (AtoOCalculator.cs:63-160), plus the
        level's normal field; then Z, rho_a and phase per site (the cell centres shifted by
        (x_shift, y_shift) metres)."""
        self.build_level(level_z, x_shift, y_shift)
        return self.level_responses()

    def build_level(self, level_z: float = 0.0, x_shift: float = 0.0, y_shift: float = 0.0):
        """The E and H level tensors of the observation level (AtoOCalculator.cs:124-160), kept in
        the operator handle for level_responses()."""
        t0 = time.perf_counter()
        self.op.calculate_green_atoo(level_z, electric=True, x_shift=x_shift, y_shift=y_shift)
        self.op.calculate_green_atoo(level_z, electric=False, x_shift=x_shift, y_shift=y_shift)
        self.atoo_green_seconds += time.perf_counter() - t0          # AtoOGreenCalc, summed over levels
        self.level_z = level_z

    def level_responses(self, chi=None) -> dict:
        """E, H, Z, rho_a and phase at the level of the latest build_level(), from chi (both
        polarisations; the latest solution by default)."""
        torch = self.torch
        op = self.op
        chi = self.chi if chi is None else chi
        normals = level_normal_fields(self.model.section, self.omega, self.level_z)
        E, H, Ea, Ha = [], [], [], []
        for pol in (0, 1):
            t0 = time.perf_counter()
            js = torch.from_numpy(self.js[pol]).to("cuda")
            jq = op.jq(chi[pol], js)
            ea = op.level_field(jq, electric=True).cpu().numpy()
            ha = op.level_field(jq, electric=False).cpu().numpy()
            self.atoo_field_seconds[pol] += time.perf_counter() - t0  # AtoOFields of this polarisation
            Ea.append(ea)
            Ha.append(ha)
            E.append(normals[pol][0][None, None, :] + ea)
            H.append(normals[pol][1][None, None, :] + ha)
        z = impedance(E[0], E[1], H[0], H[1])
        rho, phase = apparent_resistivity_phase(z, self.omega)
        return {"E": E, "H": H, "Z": z, "rho": rho, "phase": phase, "E_anomalous": Ea, "H_anomalous": Ha,
                "E_normal": [normals[p][0] for p in (0, 1)], "H_normal": [normals[p][1] for p in (0, 1)]}

    def solve_mixed(self, inner_floor=1e-6, max_refinements=10, vectors="fp32"):
        """The same two solves in mixed precision (needs precision="mixed"): chi <- chi0, then FP64
        refinement around complex64 FGMRES (complex64 Krylov vectors unless vectors="fp64") until the
        FP64 true residual reaches the tolerance."""
        self.chi.copy_(self.chi0)
        return self.solver.solve_refined_batch([self.chi0[0], self.chi0[1]], [self.chi[0], self.chi[1]],
                                               inner_floor=inner_floor, max_refinements=max_refinements,
                                               vectors=vectors)


class ModelGroupSolve:
    """Several conductivity models on one grid, background and frequency solved together: one Green
    tensor, chi0 of every (model, polarisation) from one Chi0 apply, and one batched FGMRES whose
    every apply streams the Green tensor once for all right-hand sides (the reference solves each
    model and polarisation in turn, Mt3DForwardSolver.cs:78-129, reusing the tensor through
    TensorCache, Forward/TensorCache.cs:8-35). precision="mixed" keeps a complex64 copy of the
    tensor for solve_mixed."""

    def __init__(self, models, frequency: float, settings: X.ForwardSettings | None = None,
                 precision: str = "fp64"):
        import torch
        self.torch = torch
        base = models[0]
        for m in models[1:]:
            same = (m.nx, m.ny, m.nz, m.dx, m.dy) == (base.nx, base.ny, base.nz, base.dx, base.dy) and \
                np.array_equal(m.depth, base.depth) and np.array_equal(m.thick, base.thick) and \
                np.array_equal(m.section.sigma, base.section.sigma) and \
                np.array_equal(m.section.thickness, base.section.thickness)
            if not same:
                raise ValueError("models of a group must share the grid and the background")
        self.models = list(models)
        self.frequency = frequency
        self.omega = cm.omega_of(frequency)
        self.settings = settings or X.ForwardSettings()
        K = len(models)
        self.op = X.ConvolutionOperator(base, self.omega, self.settings, precision=precision, max_rhs=2 * K)
        self.op.calculate_green_tensor()
        e0 = normal_field(base, self.omega)
        self.rhs_model = [k for k in range(K) for _ in (0, 1)]
        self.js = [scattered_current(m, self.omega, pol, e0) for m in models for pol in (0, 1)]
        zetas = [m.zeta(self.omega) for m in models]
        self.op.prepare_models("Chi0", zetas)
        js_dev = [torch.from_numpy(j).to("cuda") for j in self.js]
        self.chi0 = self.op.apply_models(js_dev, self.rhs_model)
        self.op.prepare_models("A", zetas)
        self.solver = X.AnomalyCurrentFgmresSolver(self.op, self.settings)
        self.chi = [c.clone() for c in self.chi0]

    def solve(self):
        for c, c0 in zip(self.chi, self.chi0):
            c.copy_(c0)
        self.solver.history.clear()
        return self.solver.solve_models_batch(self.chi0, self.chi, self.rhs_model)

    def solve_mixed(self, inner_floor=1e-6, max_refinements=10, vectors="fp32"):
        """solve() in mixed precision (needs precision="mixed"): FP64 refinement around complex64
        FGMRES corrections of all right-hand sides side by side, each with its own model's R, until
        every FP64 true residual reaches the tolerance."""
        for c, c0 in zip(self.chi, self.chi0):
            c.copy_(c0)
        return self.solver.solve_refined_batch(self.chi0, self.chi, inner_floor=inner_floor,
                                               max_refinements=max_refinements, vectors=vectors,
                                               models=self.rhs_model)
