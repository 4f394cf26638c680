"""Writes the golden fixtures of tests/golden/ from the oracle (the FP64 C restatement).

The reference ships no test vectors for its solver path, so these fixtures freeze the oracle's
outputs on small seeded cases: the AtoA Green tensor (after mirroring and FFT), the A and Chi0
operator applies, an FGMRES solve with its iteration history, plane-wave fields, and the
observation-level (AtoO) tensors, jQ and anomalous fields of that solve. The CPU
suite checks that the oracle still reproduces them; the GPU suite checks the HIP library
against them. Regenerate with:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from extreme_amd import cartesian_model as cm  # noqa: E402
from oracle import oracle as O  # noqa: E402

FIXTURE = os.path.join(HERE, "oracle_small.npz")


def golden_model():
    """4 x 4 x 3 cells over a three-layer background with a conductive and a resistive block."""
    sec = cm.Section1D(0.0, [0.0, 400.0, 900.0, 0.0], [0.0, 1e-2, 0.2, 2e-3])
    depth = np.array([50.0, 150.0, 400.0])
    thick = np.array([100.0, 250.0, 300.0])
    sigma = np.empty((4, 4, 3))
    for k, bg in enumerate([1e-2, 1e-2, 0.2]):
        sigma[:, :, k] = bg
    sigma[1:3, 1:3, 0] = 1.0
    sigma[0:2, 2:4, 2] = 1e-3
    sigma[3, 0, 1] = 0.5
    return cm.CartesianModel(4, 4, 300.0, 200.0, depth, thick, sigma, sec, "golden", [0.5])


def compute():
    m = golden_model()
    w = cm.omega_of(m.frequencies[0])
    G = O.green_atoa(m, w)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op = O.Operator(m, G)
    rng = np.random.default_rng(20261202)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    op.prepare("A", z, z0)
    y_a = op.apply(x)
    op.prepare("Chi0", z, z0)
    y_chi0 = op.apply(x)
    e0 = O.normal_field(m, w)
    js = O.jscattered(m, 0, z, z0, e0)
    chi0 = op.apply(js)
    op.prepare("A", z, z0)
    chi, info, hist = O.fgmres(op.raw_apply, chi0, chi0, 30, 10, 100, 1e-8)
    depths = np.array([0.0, 75.0, 399.0, 400.0, 1000.0, 1299.0])
    freqs = np.array([1e-3, 0.5, 30.0])
    pw_e = np.array([[O.plane_wave_e(m.section, cm.omega_of(f), d) for d in depths] for f in freqs])
    pw_h = np.array([[O.plane_wave_h(m.section, cm.omega_of(f), d) for d in depths] for f in freqs])
    # observation level at the surface: level tensors, jQ of the solved chi, anomalous fields
    Ge = O.green_atoo_level(m, w, 0.0, True)
    Gh = O.green_atoo_level(m, w, 0.0, False)
    _, jq = O.escattered_jq(m, z, z0, chi, js)
    field_e = O.atoo_field(m, Ge, True, jq)
    field_h = O.atoo_field(m, Gh, False, jq)
    out = {f"green_{c}": G[c] for c in O.COMPONENTS}
    out.update({f"atoo_e_{c}": v for c, v in Ge.items()})
    out.update({f"atoo_h_{c}": v for c, v in Gh.items()})
    out.update(jq=jq, level_field_e=field_e, level_field_h=field_h)
    out.update(x=x, y_a=y_a, y_chi0=y_chi0, js=js, chi0=chi0, chi=chi,
               fgmres_iterations=np.int64(info["iterations"]), fgmres_matvecs=np.int64(info["matvecs"]),
               fgmres_bea=np.array([h[2] for h in hist if h[0] == 1]), pw_depths=depths, pw_freqs=freqs, pw_e=pw_e, pw_h=pw_h)
    return out


def load():
    with np.load(FIXTURE, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


if __name__ == "__main__":
    np.savez_compressed(FIXTURE, **compute())
    print("wrote", FIXTURE, os.path.getsize(FIXTURE), "bytes")
