"""The reference's COMMEMI 3D-2 example run end to end on the GPU from its own model files:
project -> model reader -> Green tensor, FGMRES (both polarisations) -> observation level -> the
reference's result tables. The reference ships no result files for it, so the checks are the
model's own symmetry and the layered-earth limit away from the blocks."""
import os

import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import forward as F
from extreme_amd import project_io as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "golden", "commemi3d2", "commemi.xmodel")


def test_commemi_end_to_end(torch_cuda, tmp_path):
    proj = tmp_path / "forward.xproj"
    proj.write_text(f"""<?xml version="1.0" encoding="utf-8"?>
<ExtremeProject version="0.7.2">
  <ResultsPath>results</ResultsPath>
  <ModelFile>{MODEL}</ModelFile>
  <Frequencies><F>0.1</F></Frequencies>
  <Observations><Tablets><T z="0"></T></Tablets></Observations>
  <Settings><Forward><Residual>1E-10</Residual><InnerBufferLength>10</InnerBufferLength>
    <OuterBufferLength>15</OuterBufferLength><MaxRepeatsNumber>100</MaxRepeatsNumber>
    <NumberOfHankels>5</NumberOfHankels></Forward></Settings>
</ExtremeProject>""")
    logs = []
    files = P.run_project(str(proj), log=logs.append)
    print(logs, flush=True)
    assert [os.path.basename(f) for f in files] == ["fields_freq0000000000.1000000000",
                                                    "responses_freq0000000000.1000000000"]
    names, rows = P.read_result_table(files[1])
    assert rows.shape == (1600, 28) and np.all(np.isfinite(rows))
    z = (rows[:, 4:12:2] + 1j * rows[:, 5:12:2]).reshape(40, 40, 4)   # Zxx, Zxy, Zyx, Zyy per (i, j)
    # the model is symmetric under y -> -y about the block centre (j -> 39 - j): Zxy and Zyx even,
    # Zxx and Zyy odd (written with 7 significant digits)
    zs = z[:, ::-1, :]
    scale = np.abs(z[..., 1]).max()
    even = max(np.abs(z[..., 1] - zs[..., 1]).max(), np.abs(z[..., 2] - zs[..., 2]).max()) / scale
    odd = max(np.abs(z[..., 0] + zs[..., 0]).max(), np.abs(z[..., 3] + zs[..., 3]).max()) / scale
    print(f"COMMEMI y-mirror: even parts {even:.2e}, odd parts {odd:.2e}", flush=True)
    assert even < 1e-5 and odd < 1e-5
    # far from the blocks (the grid's corners, >= 27 km from them, skin depth 5 km) the response
    # approaches the 1-D layered earth's
    m = P.read_model(MODEL)
    w = cm.omega_of(0.1)
    (e1, h1), _ = F.level_normal_fields(m.section, w, 0.0)
    z1d = e1[0] / h1[1]
    rho1d = abs(z1d) ** 2 / (w * cm.MU0)
    rho = rows[:, 12:16].reshape(40, 40, 4)
    corners = [rho[0, 0], rho[0, 39], rho[39, 0], rho[39, 39]]
    dev = max(max(abs(c[1] / rho1d - 1), abs(c[2] / rho1d - 1)) for c in corners)
    print(f"COMMEMI corner rho_xy, rho_yx vs 1-D {rho1d:.4f} ohm-m: max deviation {dev:.3%}", flush=True)
    assert dev < 5e-3                               # measured 0.045 %
    # the fields table carries the blocks' anomalous fields (the last 24 columns)
    names_f, fields = P.read_result_table(files[0])
    assert fields.shape == (1600, 52) and np.abs(fields[:, 28:]).max() > 0
    # the frequency's timings, appended to common.dat beside the results folder
    rows = (tmp_path / "common.dat").read_text().splitlines()
    assert len(rows) == 2 and rows[0].startswith("Date           Time        nx")
    cols = rows[1].split()
    assert cols[2:8] == ["40", "40", "40", "5", "1", "1"]
    times = [float(c) for c in cols[8:]]
    print(f"common.dat: {rows[1]}", flush=True)
    assert len(times) == 6 and all(t > 0 for t in times)
    # the frequency's event breakdown beside the result tables (ExtremeMtSolverRunner.cs:135-137)
    prof = (tmp_path / "results" / "profiling_mpi0001_threads1_freq0.1000.txt").read_text()
    print(prof, flush=True)
    lines = prof.splitlines()
    assert lines[2] == "nx = 40, ny = 40, nz = 40"
    mult = int(next(l for l in lines if l.startswith("Total number of multiplications:")).split(":")[1])
    assert mult > 0
    for ev in ("GreenAtoATotal:", "SolveCie:", "ObservationsFullCalculation:", "CalcChi0:"):
        assert ev in lines
    stage = [l.split() for l in lines if l.strip().startswith("OperatorA")]
    assert [s[0] for s in stage] == ["OperatorAForwardFft", "OperatorAMultiplication", "OperatorABackwardFft"]
    assert all(int(s[4]) == mult for s in stage)
    assert "Total Covered:" in lines[-1]


def test_commemi_reference_project(torch_cuda, tmp_path):
    """The example's own project: four tablets shifted by (+-20 km, +-20 km), three frequencies.
    Where tablets overlap they observe the same sites with the same scalar plan, so their responses
    agree to the written digits; tablets (x, -20 km) and (x, +20 km) are y-mirror images."""
    logs = []
    files = P.run_project(os.path.join(HERE, "golden", "commemi3d2", "forward.xproj"),
                          results_path=str(tmp_path / "results"), log=logs.append)
    print(logs, flush=True)
    assert len(files) == 6
    assert len((tmp_path / "common.dat").read_text().splitlines()) == 1 + 3     # header + 3 frequencies
    proj = P.read_project(os.path.join(HERE, "golden", "commemi3d2", "forward.xproj"))
    shifts = [(l.x_shift, l.y_shift) for l in proj.levels]
    for f in files[1::2]:
        _, rows = P.read_result_table(f)
        assert rows.shape == (4 * 1600, 28) and np.all(np.isfinite(rows))
        tab = {s: rows[k * 1600:(k + 1) * 1600].reshape(40, 40, 28) for k, s in enumerate(shifts)}
        a, b = tab[(-20000.0, -20000.0)], tab[(20000.0, -20000.0)]
        assert np.array_equal(a[20:, :, 0:3], b[:20, :, 0:3])              # same site coordinates
        za, zb = a[20:, :, 4:12], b[:20, :, 4:12]
        overlap = np.abs(za - zb).max() / np.abs(za).max()
        c, d = tab[(-20000.0, -20000.0)], tab[(-20000.0, 20000.0)]
        z1 = c[..., 4:12:2] + 1j * c[..., 5:12:2]
        z2 = (d[..., 4:12:2] + 1j * d[..., 5:12:2])[:, ::-1]
        scale = np.abs(z1[..., 1]).max()
        even = max(np.abs(z1[..., k] - z2[..., k]).max() for k in (1, 2)) / scale
        odd = max(np.abs(z1[..., k] + z2[..., k]).max() for k in (0, 3)) / scale
        print(f"{os.path.basename(f)}: overlap {overlap:.2e}, y-mirror even {even:.2e} odd {odd:.2e}", flush=True)
        assert overlap < 1e-5 and even < 1e-5 and odd < 1e-5
    # 0.1 Hz, the far corner of tablet (-20 km, -20 km) (>= 39 km from the blocks): the 1-D earth
    _, rows = P.read_result_table(files[1])
    m = P.read_model(MODEL)
    w = cm.omega_of(0.1)
    (e1, h1), _ = F.level_normal_fields(m.section, w, 0.0)
    rho1d = abs(e1[0] / h1[1]) ** 2 / (w * cm.MU0)
    corner = rows[0, 12:16]
    dev = max(abs(corner[1] / rho1d - 1), abs(corner[2] / rho1d - 1))
    print(f"far corner rho_xy, rho_yx vs 1-D: {dev:.3%}", flush=True)
    assert dev < 5e-3
