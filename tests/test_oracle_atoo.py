"""Pins the oracle's AtoO restatement (observation-level tensors and fields) without the reference:
the thin-layer limit against the independently restated AtoA tensor, the lateral parities of the
explicitly computed negative offsets, and the FFT convolution against a direct sum. CPU only."""
import numpy as np
import pytest

from conftest import rel
from extreme_amd import cartesian_model as cm


def _thin_model(dz, nz=12, nx=4, ny=4):
    sec = cm.Section1D(0.0, [0.0, 2000.0, 0.0], [0.0, 1e-2, 1e-3])
    depth = 500.0 + np.arange(nz) * dz
    thick = np.full(nz, dz)
    return cm.CartesianModel(nx, ny, 200.0, 200.0, depth, thick, np.full((nx, ny, nz), 1e-2), sec, "thin", [1.0])


def _atoa_row(oracle, G, m, comp, rc, tr):
    L, nz = 4 * m.nx * m.ny, m.nz
    if comp in ("xz", "yz"):
        return G[comp].reshape(L, nz, nz)[:, rc, tr]
    a, b = min(rc, tr), max(rc, tr)
    return G[comp].reshape(L, -1)[:, a + b * (b + 1) // 2]


def test_atoo_electric_is_the_thin_layer_limit_of_atoa(oracle):
    """The AtoA row of a receiver layer of thickness dz is the volume-to-point tensor integrated over
    the layer: AtoA / (dz AtoO(layer centre)) = 1 + O(dz^2), for every component; zx, zy are the
    reciprocity images -xz^T, -yz^T of the AtoA tensor."""
    w = cm.omega_of(1.0)
    errs = []
    for dz in (5.0, 2.5):
        m = _thin_model(dz)
        Ga = oracle.green_atoa(m, w, fft=False)
        rc = 6
        Go = oracle.green_atoo_level(m, w, m.depth[rc] + dz / 2, True, fft=False)
        worst = 0.0
        for tr in (0, 2, 11):
            pairs = [(c, _atoa_row(oracle, Ga, m, c, rc, tr), Go[c][:, tr]) for c in ("xx", "yy", "zz", "xy", "xz", "yz")]
            pairs.append(("zx", -_atoa_row(oracle, Ga, m, "xz", tr, rc), Go["zx"][:, tr]))
            pairs.append(("zy", -_atoa_row(oracle, Ga, m, "yz", tr, rc), Go["zy"][:, tr]))
            for name, a, o in pairs:
                e = np.linalg.norm(a - dz * o) / np.linalg.norm(dz * o)
                worst = max(worst, e)
                assert e < 2e-3, (dz, tr, name, e)
        errs.append(worst)
    # second order: halving dz divides the error by about four
    assert 3.0 < errs[0] / errs[1] < 5.0, errs


# lateral parities under x -> -x and y -> -y: E is a polar vector, H an axial one, J polar
_S_POLAR_X, _S_POLAR_Y = (-1, 1, 1), (1, -1, 1)
_S_AXIAL_X, _S_AXIAL_Y = (1, -1, -1), (-1, 1, -1)
_IDX = {"x": 0, "y": 1, "z": 2}


@pytest.mark.parametrize("electric", [True, False])
def test_atoo_lateral_parities(oracle, electric):
    """Negative offsets are integrated explicitly (with the 'Symm' sign tables), not mirrored:
    G(-x, y) and G(x, -y) must carry the parities of the field and the current."""
    m = cm.small_model(5, 4, 3, seed=8)
    w = cm.omega_of(0.3)
    G = oracle.green_atoo_level(m, w, 0.0, electric, fft=False)
    nx, ny = m.nx, m.ny
    fx, fy = (_S_POLAR_X, _S_POLAR_Y) if electric else (_S_AXIAL_X, _S_AXIAL_Y)
    for comp, g in G.items():
        a, b = _IDX[comp[0]], _IDX[comp[1]]
        px, py = fx[a] * _S_POLAR_X[b], fy[a] * _S_POLAR_Y[b]
        g = g.reshape(2 * nx, 2 * ny, -1)
        scale = np.abs(g).max()
        if scale == 0:
            continue
        for i in range(1, nx):
            for j in range(0, ny):
                assert np.allclose(g[2 * nx - i, j], px * g[i, j], rtol=1e-9, atol=1e-12 * scale), (comp, i, j)
        for i in range(0, nx):
            for j in range(1, ny):
                assert np.allclose(g[i, 2 * ny - j], py * g[i, j], rtol=1e-9, atol=1e-12 * scale), (comp, i, j)
        assert np.all(g[nx] == 0) and np.all(g[:, ny] == 0)


def _direct_field(m, Gs, electric, jq):
    nx, ny, nz = m.nx, m.ny, m.nz
    jq = jq.reshape(nx, ny, 3, nz)
    g = {c: v.reshape(2 * nx, 2 * ny, nz) for c, v in Gs.items()}
    out = np.zeros((nx, ny, 3), dtype=complex)
    for i in range(nx):
        for j in range(ny):
            for i2 in range(nx):
                for j2 in range(ny):
                    di, dj = (i - i2) % (2 * nx), (j - j2) % (2 * ny)
                    sx, sy, sz = jq[i2, j2, 0], jq[i2, j2, 1], jq[i2, j2, 2]
                    G_ = {c: v[di, dj] for c, v in g.items()}
                    if electric:
                        out[i, j, 0] += G_["xx"] @ sx + G_["xy"] @ sy + G_["xz"] @ sz
                        out[i, j, 1] += G_["xy"] @ sx + G_["yy"] @ sy + G_["yz"] @ sz
                        out[i, j, 2] += G_["zx"] @ sx + G_["zy"] @ sy + G_["zz"] @ sz
                    else:
                        out[i, j, 0] += -(G_["xx"] @ sx + G_["xy"] @ sy + G_["xz"] @ sz)
                        out[i, j, 1] += G_["yx"] @ sx + G_["xx"] @ sy + G_["yz"] @ sz
                        out[i, j, 2] += -(G_["zx"] @ sx) + G_["zy"] @ sy
    return out


@pytest.mark.parametrize("electric", [True, False])
def test_atoo_field_matches_direct_sum(oracle, electric):
    """FFT convolution with zero padding, 1/(4 Nx Ny) and cropping = direct sum over source cells."""
    m = cm.small_model(4, 3, 2, seed=12)
    w = cm.omega_of(2.0)
    Gs = oracle.green_atoo_level(m, w, 0.0, electric, fft=False)
    Gf = oracle.green_atoo_level(m, w, 0.0, electric, fft=True)
    rng = np.random.default_rng(4)
    jq = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    assert rel(oracle.atoo_field(m, Gf, electric, jq), _direct_field(m, Gs, electric, jq)) < 1e-12


def test_impedance_of_the_normal_field_is_the_1d_impedance(oracle):
    """With no anomaly the level fields are the plane wave: Zxy = E0/H0, Zyx = -E0/H0, Zxx = Zyy = 0
    (Mt3DForwardSolver.CalculateNormalFieldE/H(level) :236-271 and the 2x2 tensor)."""
    sec = cm.Section1D(0.0, [0.0, 1000.0, 0.0], [0.0, 1e-2, 1e-1])
    w = cm.omega_of(0.1)
    e0, h0 = oracle.plane_wave_e(sec, w, 0.0), oracle.plane_wave_h(sec, w, 0.0)
    e1, e2 = np.array([e0, 0, 0]), np.array([0, e0, 0])
    h1, h2 = np.array([0, h0, 0]), np.array([-h0, 0, 0])
    z = oracle.impedance(e1, e2, h1, h2)
    assert z[0, 0] == 0 and z[1, 1] == 0
    assert abs(z[0, 1] - e0 / h0) < 1e-15 * abs(e0 / h0)
    assert abs(z[1, 0] + e0 / h0) < 1e-15 * abs(e0 / h0)
    rho, phi = oracle.rho_phase(z, w)
    assert abs(rho[0, 1] - abs(e0 / h0) ** 2 / (w * cm.MU0)) < 1e-12 * rho[0, 1]
