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


# --- shifted observation levels ---------------------------------------------------------------

def test_level_rho_range_closed_forms(oracle):
    """FindRhoMin / FindRhoMax of a level (ScalarPlansCreater.cs:180-266) against hand-worked values."""
    m = cm.small_model(6, 5, 2, seed=1)                   # dx = dy = 200
    dx, dy, nx, ny = m.dx, m.dy, m.nx, m.ny
    diag = np.hypot(nx * dx - dx / 2, ny * dy - dy / 2)
    assert oracle.atoo_level_rho_range(m) == (100.0, diag, False)                 # the AtoA range
    # two cells right, one down: footprint x in [300, 1500], y in [-300, 700]; the farthest corner
    # from the anomaly's corner centres is (-300 - 0, ...) -> (1500, -300) from (0, 800)
    rmin, rmax, zero = oracle.atoo_level_rho_range(m, 2 * dx, -dy)
    assert (rmin, zero) == (100.0, False) and rmax == np.hypot(1500.0, 1100.0)
    # 0.3 of a cell: receivers 60 m from a cell centre, 40 m from the nearer cell edge
    assert oracle.atoo_level_rho_range(m, 60.0, 0.0)[0::2] == (40.0, False)
    # half a cell: receivers on cell edges -> zero rho, rho_min = 0.001 min(dx, dy)
    assert oracle.atoo_level_rho_range(m, dx / 2, 0.0)[0::2] == (0.2, True)
    # no overlap in either axis: the distances from the footprint edges to the nearest cell centres
    rmin, _, zero = oracle.atoo_level_rho_range(m, 3 * nx * dx, 3 * ny * dy)
    assert (rmin, zero) == (min((2 * nx + 0.5) * dx, (2 * ny + 0.5) * dy), False)


@pytest.mark.parametrize("electric", [True, False])
def test_whole_cell_shift_moves_the_tensor(oracle, electric):
    """A level shifted by (2 dx, -dy) sees the offsets of the unshifted level moved by (2, -1) cells:
    with the same scalar plan the two tensors agree entry by entry where both offsets exist."""
    m = cm.small_model(6, 5, 3, seed=3)
    w = cm.omega_of(0.5)
    s, t = 2, -1
    rr = oracle.atoo_level_rho_range(m, s * m.dx, t * m.dy)
    Gs = oracle.green_atoo_level(m, w, 0.0, electric, fft=False, x_shift=s * m.dx, y_shift=t * m.dy)
    G0 = oracle.green_atoo_level(m, w, 0.0, electric, fft=False, rho_range=rr)
    nx, ny = m.nx, m.ny
    n = 0
    for c in Gs:
        a, b = Gs[c].reshape(2 * nx, 2 * ny, -1), G0[c].reshape(2 * nx, 2 * ny, -1)
        scale = np.abs(b).max()
        for i in range(-nx + 1, nx):
            for j in range(-ny + 1, ny):
                if abs(i + s) < nx and abs(j + t) < ny:
                    assert np.abs(a[i % (2 * nx), j % (2 * ny)] - b[(i + s) % (2 * nx), (j + t) % (2 * ny)]).max() \
                        <= 1e-13 * scale, (c, i, j)
                    n += 1
    assert n > 0


@pytest.mark.parametrize("frac", [0.3, 0.5])
@pytest.mark.parametrize("electric", [True, False])
def test_mirror_shifts(oracle, electric, frac):
    """Levels shifted by +f dx and -f dx are mirror images under x -> -x: G+(i, j) = p G-(-i, j) with
    the lateral parity p of the component (f = 0.5 puts the receivers on cell edges: zero-rho plan)."""
    m = cm.small_model(5, 4, 2, seed=6)
    w = cm.omega_of(0.7)
    sh = frac * m.dx
    assert oracle.atoo_level_rho_range(m, sh, 0.0) == oracle.atoo_level_rho_range(m, -sh, 0.0)
    Gp = oracle.green_atoo_level(m, w, 0.0, electric, fft=False, x_shift=sh)
    Gm = oracle.green_atoo_level(m, w, 0.0, electric, fft=False, x_shift=-sh)
    fx = _S_POLAR_X if electric else _S_AXIAL_X
    nx, ny = m.nx, m.ny
    for c in Gp:
        p = fx[_IDX[c[0]]] * _S_POLAR_X[_IDX[c[1]]]
        a, b = Gp[c].reshape(2 * nx, 2 * ny, -1), Gm[c].reshape(2 * nx, 2 * ny, -1)
        assert np.all(np.isfinite(a))
        scale = np.abs(a).max()
        for i in range(-nx + 1, nx):
            assert np.abs(a[i % (2 * nx)] - p * b[(-i) % (2 * nx)]).max() <= 1e-12 * scale, (c, i)


@pytest.mark.parametrize("electric", [True, False])
def test_zero_rho_point(oracle, electric):
    """CalculateZeroRhoForE/H (ScalarMathUtils.cs:33-81): the merged grid gains rho = 0 in front and
    nothing else changes; I1, I2 (I5) are 0 there; I3, I4 take the value of the a + b rho^2 through the
    first item's 2nd and 3rd radii (rho_min a and rho_min a^2, a the rho step)."""
    m = cm.small_model(5, 4, 2, seed=6)
    w = cm.omega_of(0.7)
    rmin, rmax, zero = oracle.atoo_level_rho_range(m, m.dx / 2, 0.0)
    assert zero
    rho, I = oracle.atoo_level_scalars(m, w, 0.0, (rmin, rmax, True), electric)
    rho_n, I_n = oracle.atoo_level_scalars(m, w, 0.0, (rmin, rmax, False), electric)
    assert rho[0] == 0 and np.array_equal(rho[1:], rho_n) and np.array_equal(I[:, :, 1:], I_n)
    a = oracle.scalar_plan(m, rho_range=(rmin, rmax, True))["rho"][0][1:3]
    assert a[0] == rmin * oracle.hankel_rho_step() and a[1] == a[0] * oracle.hankel_rho_step()
    k1, k2 = np.searchsorted(rho, a[0]), np.searchsorted(rho, a[1])
    assert rho[k1] == a[0] and rho[k2] == a[1]
    a2 = oracle.hankel_rho_step() ** 2
    for c in range(I.shape[0]):
        if c in (2, 3):
            # the extrapolation is exact for a + b rho^2: check it on that function first
            q = lambda r: 3.0 - 2.0j + (0.5 + 1j) * r * r        # noqa: E731
            assert abs((q(a[0]) * a2 - q(a[1])) / (a2 - 1) - q(0.0)) < 1e-12
            want = (I[c, :, k1] * a2 - I[c, :, k2]) / (a2 - 1)
            assert np.allclose(I[c, :, 0], want, rtol=1e-14, atol=0)
        else:
            assert np.all(I[c, :, 0] == 0)
