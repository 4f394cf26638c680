"""Pins the oracle (the FP64 C restatement) against what mathematics fixes, independently of its
own formulas: textbook DFTs, known Hankel transform pairs, a dense restatement of the packed
contraction, a direct (FFT-free) spatial convolution, a dense LU solve of the operator equation,
the zrotg identities, the homogeneous half-space and the layered impedance recursion, and
translation invariance of the Green tensor in a homogeneous medium.

Each test is chosen so that a plausible slip (a transposed operand, a dropped sign, an index off
by one, a wrong normalisation) fails it. CPU only.
"""
import numpy as np
import pytest

from conftest import rel
from extreme_amd import cartesian_model as cm


# ---------------------------------------------------------------------------- DFT (A4, A6, G13)

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 16, 24, 25, 48, 50, 64, 80, 96, 150, 256, 384, 400])
def test_fft1_matches_textbook_dft(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x       # sign -1 forward, unnormalised
    assert rel(oracle.fft1(x, -1), dft) < 1e-13
    assert rel(oracle.fft1(x, +1), np.conj(np.exp(-2j * np.pi * np.outer(k, k) / n)) @ x) < 1e-13
    assert rel(oracle.fft1(x, -1), np.fft.fft(x)) < 1e-13


@pytest.mark.parametrize("n1,n2,nc", [(8, 6, 5), (16, 16, 3), (4, 24, 2), (10, 80, 2), (50, 12, 3)])
def test_fft2_batch_matches_numpy(oracle, n1, n2, nc):
    rng = np.random.default_rng(n1 * 100 + n2)
    x = rng.standard_normal((n1, n2, nc)) + 1j * rng.standard_normal((n1, n2, nc))
    got = oracle.fft2_batch(x.copy(), n1, n2, nc, -1)
    assert rel(got, np.fft.fft2(x, axes=(0, 1))) < 1e-13
    back = oracle.fft2_batch(got.copy(), n1, n2, nc, +1)
    assert rel(back, x * n1 * n2) < 1e-13                       # unnormalised inverse


# ---------------------------------------------------------------------------- Hankel filter (G3)

def _pairs(lam, rho):
    """Known transforms: (kernel samples, order, exact integral of kernel*J_order(lambda rho))."""
    out = []
    for z in (0.3, 1.0, 5.0):
        out.append((np.exp(-lam * z), 1, (1 - z / np.sqrt(rho ** 2 + z ** 2)) / rho, f"exp{z}-J1", 1e-12))
        out.append((np.exp(-lam * z), 0, 1 / np.sqrt(rho ** 2 + z ** 2), f"exp{z}-J0", 5e-7))
    for a in (0.5, 2.0):
        out.append((lam * np.exp(-a * lam ** 2), 0, np.exp(-rho ** 2 / (4 * a)) / (2 * a), f"gauss{a}-J0", 1e-12))
    return out


def test_hankel_filter_exact_grid_transform_pairs(oracle):
    """With the exact 10^(1/40) spacing the 431-tap filters reproduce closed-form transforms:
    pins the filter data (both tables) and the reversed-tap indexing of ConvoluteWithFast."""
    h0, h1 = oracle.hankel_table()
    assert h0.shape == (431,) and h1.shape == (431,)
    step = 10 ** (1 / 40)
    nrho = 60
    rho = step ** np.arange(nrho)
    lam = step ** (130 - np.arange(431 + nrho))                 # lambda[130] = 1/rho[0]
    for fk, order, exact, name, tol in _pairs(lam, rho):
        got = oracle.hankel_filter(h1 if order else h0, fk, nrho).real / rho
        err = np.max(np.abs(got - exact)) / np.max(np.abs(exact))
        assert err < tol, (name, err)
    # a plausible slip: taps not reversed, or the grid shifted by one sample
    fk = np.exp(-lam)
    rev = oracle.hankel_filter(h1[::-1].copy(), fk, nrho).real / rho
    assert np.max(np.abs(rev - (1 - 1 / np.sqrt(rho ** 2 + 1)) / rho) * rho) > 1e-3
    sh = oracle.hankel_filter(h1, fk[1:], nrho - 1).real / rho[:-1]
    assert np.max(np.abs(sh - (1 - 1 / np.sqrt(rho[:-1] ** 2 + 1)) / rho[:-1]) * rho[:-1]) > 1e-3


def test_hankel_filter_reference_grid(oracle):
    """The reference rounds the step to float (HankelCoefficients.cs:67): the grid is geometric with
    that ratio, lambda[130] = 1/rho0, and the transforms stay within the rounding's 1e-6."""
    step = oracle.hankel_rho_step()
    assert step == float(np.float32(10 ** (1 / 40)))
    rho0, nrho = 37.5, 40
    lam = oracle.hankel_lambdas(rho0, nrho)
    assert lam.shape == (431 + nrho,)
    assert lam[130] == 1 / rho0
    assert np.allclose(lam[:-1] / lam[1:], step, rtol=1e-14, atol=0)
    rho = rho0 * step ** np.arange(nrho)
    h0, h1 = oracle.hankel_table()
    for fk, order, exact, name, _ in _pairs(lam * rho0, rho / rho0):
        # same transforms in scaled variables: integral f(lam rho0) J(lam rho) dlam = F(rho/rho0)/rho0
        got = oracle.hankel_filter(h1 if order else h0, fk, nrho).real / rho
        err = np.max(np.abs(got - exact / rho0)) / np.max(np.abs(exact / rho0))
        assert err < 2e-6, (name, err)


# ---------------------------------------------------------------------------- contraction (A5)

def _dense_blocks(green, nz, l):
    """3nz x 3nz matrix of wavenumber l from the packed reference layout: xx, xy, yy, zz packed
    upper (element (rc,tr), rc <= tr, at rc + tr(tr+1)/2, symmetric in rc/tr), xz, yz full rc-major;
    the z-row is the reciprocity image -xz^T, -yz^T (TranceiverUtils.cs:10-16,55-74)."""
    def sym(c):
        m = np.zeros((nz, nz), dtype=complex)
        for tr in range(nz):
            for rc in range(tr + 1):
                m[rc, tr] = m[tr, rc] = green[c][l, rc + tr * (tr + 1) // 2]
        return m
    xz = green["xz"][l].reshape(nz, nz)
    yz = green["yz"][l].reshape(nz, nz)
    xx, xy, yy, zz = sym("xx"), sym("xy"), sym("yy"), sym("zz")
    return np.block([[xx, xy, xz], [xy, yy, yz], [-xz.T, -yz.T, zz]])


@pytest.mark.parametrize("nz", [1, 2, 3, 5])
def test_contract_matches_dense_blocks(oracle, nz):
    rng = np.random.default_rng(nz)
    L = 7
    S = nz * (nz + 1) // 2
    green = {c: rng.standard_normal((L, nz * nz if c in ("xz", "yz") else S))
             + 1j * rng.standard_normal((L, nz * nz if c in ("xz", "yz") else S)) for c in oracle.COMPONENTS}
    u = rng.standard_normal((L, 3 * nz)) + 1j * rng.standard_normal((L, 3 * nz))
    got = oracle.contract(nz, green, u).reshape(L, 3 * nz)
    want = np.stack([_dense_blocks(green, nz, l) @ u[l] for l in range(L)])
    assert rel(got, want) < 1e-14


# ---------------------------------------------------------------------------- operator (A1-A9)

# lateral parities of the layered-medium Green tensor under x -> -x and y -> -y
_PARITY = {"xx": (1, 1), "xy": (-1, -1), "xz": (-1, 1), "yy": (1, 1), "yz": (1, -1), "zz": (1, 1)}


def _brute_force_apply(model, green_spatial, zeta, zeta0, x, kind):
    """A x = x - R x - b conv(G, f R x) (Chi0: R0 x + b conv) with conv a direct double sum over
    source cells; G at negative lateral offsets from the physical parities of the first quadrant."""
    nx, ny, nz = model.nx, model.ny, model.nz
    x = x.reshape(nx, ny, 3, nz)
    z = zeta.reshape(nx, ny, nz)
    if kind == "A":
        r = (z - zeta0) / (z + np.conj(zeta0))
    else:
        r = np.sqrt(zeta0.real + 0j) / (z + np.conj(zeta0))
    f = np.sqrt(zeta0.real)
    b = 2 * np.sqrt(zeta0.real) / model.thick
    rx = r[:, :, None, :] * x
    u = f * rx
    quad = {c: green_spatial[c].reshape(2 * nx, 2 * ny, -1)[:nx, :ny] for c in green_spatial}
    conv = np.zeros_like(x)
    for di in range(-nx + 1, nx):
        for dj in range(-ny + 1, ny):
            g = {}
            for c, (px, py) in _PARITY.items():
                g[c] = quad[c][abs(di), abs(dj)] * (px if di < 0 else 1) * (py if dj < 0 else 1)
            dense = _dense_blocks({c: g[c][None] for c in g}, nz, 0)
            i0, i1 = max(0, di), min(nx, nx + di)
            j0, j1 = max(0, dj), min(ny, ny + dj)
            src = u[i0 - di:i1 - di, j0 - dj:j1 - dj].reshape(i1 - i0, j1 - j0, 3 * nz)
            conv[i0:i1, j0:j1] += (src @ dense.T).reshape(i1 - i0, j1 - j0, 3, nz)
    y = b * conv
    return (x - y - rx).reshape(-1) if kind == "A" else (y + rx).reshape(-1)


@pytest.mark.parametrize("kind", ["A", "Chi0"])
def test_operator_apply_matches_direct_convolution(oracle, kind):
    m = cm.small_model(4, 3, 2, dx=250.0, dz=100.0, contrast=30.0, seed=11)
    w = cm.omega_of(1.0)
    Gs = oracle.green_atoa(m, w, fft=False)
    Gf = oracle.green_atoa(m, w, fft=True)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op = oracle.Operator(m, Gf)
    op.prepare(kind, z, z0)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    want = _brute_force_apply(m, Gs, z, z0, x, kind)
    assert rel(op.apply(x), want) < 1e-12


def test_operator_identity_when_no_anomaly(oracle):
    """zeta == zeta0 gives R = 0, so A = I exactly, whatever the Green tensor."""
    m = cm.small_model(4, 4, 3, seed=2)
    w = cm.omega_of(0.5)
    z0 = m.zeta0(w)
    z = np.broadcast_to(z0[None, None, :], (m.nx, m.ny, m.nz)).reshape(-1).copy()
    op = oracle.Operator(m, oracle.green_atoa(m, w))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    op.prepare("A", z, z0)
    assert np.array_equal(op.apply(x), x)


def test_r_bounded_by_one():
    """|zeta - zeta0|^2 = |zeta + conj zeta0|^2 - 4 Re zeta Re zeta0, so |R| <= 1 for Re >= 0."""
    rng = np.random.default_rng(3)
    z = rng.uniform(0, 10, 1000) - 1j * rng.uniform(0, 1e-6, 1000)
    z0 = rng.uniform(1e-4, 10, 1000) - 1j * rng.uniform(0, 1e-6, 1000)
    r = (z - z0) / (z + np.conj(z0))
    assert np.all(np.abs(r) <= 1 + 1e-15)
    assert np.allclose(np.abs(z - z0) ** 2, np.abs(z + np.conj(z0)) ** 2 - 4 * z.real * z0.real, rtol=1e-12)


# ---------------------------------------------------------------------------- FGMRES (K1-K13)

def test_fgmres_matches_dense_lu(oracle):
    m = cm.small_model(3, 3, 2, contrast=100.0, seed=4)
    w = cm.omega_of(3.0)
    G = oracle.green_atoa(m, w)
    op = oracle.Operator(m, G)
    op.prepare("A", m.zeta(w).reshape(-1), m.zeta0(w))
    n = m.n
    A = np.stack([op.apply(e) for e in np.eye(n, dtype=complex)], axis=1)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x_lu = np.linalg.solve(A, b)
    x, info, hist = oracle.fgmres(op.raw_apply, b, b.copy(), 30, 10, 100, 1e-12)
    cond = np.linalg.cond(A)
    assert info["converged"]
    assert rel(x, x_lu) < 1e-12 * cond * 10
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-11
    # the Arnoldi estimate never increases within a restart cycle
    beas = [h[2] for h in hist]
    assert all(beas[i + 1] <= beas[i] * (1 + 1e-12) for i in range(len(beas) - 1) if hist[i + 1][0] == hist[i][0])


def test_fgmres_restarts_reach_the_same_solution(oracle):
    """A short outer buffer forces ComputeTheTrue restarts; the solution must not change."""
    m = cm.small_model(3, 3, 2, contrast=100.0, seed=4)
    w = cm.omega_of(3.0)
    op = oracle.Operator(m, oracle.green_atoa(m, w))
    op.prepare("A", m.zeta(w).reshape(-1), m.zeta0(w))
    b = np.random.default_rng(2).standard_normal(m.n) + 0j
    x1, i1, _ = oracle.fgmres(op.raw_apply, b, b.copy(), 30, 10, 100, 1e-11)
    x2, i2, _ = oracle.fgmres(op.raw_apply, b, b.copy(), 2, 2, 200, 1e-11)
    assert i1["converged"] and i2["converged"]
    assert rel(x1, x2) < 1e-9


@pytest.mark.parametrize("a,b", [(3 + 4j, 1 - 2j), (1e-200 + 0j, 5j), (0j, 2 - 1j), (-2 + 0j, 0j), (1e150 + 1e150j, 1e150j)])
def test_zrotg_identities(oracle, a, b):
    c, s = oracle.zrotg(a, b)
    r = c * a + s * b
    assert abs(c * c + abs(s) ** 2 - 1) < 1e-15
    assert abs(-np.conj(s) * a + c * b) <= 1e-15 * max(abs(a), abs(b))
    assert abs(abs(r) - np.hypot(abs(a), abs(b))) <= 1e-15 * abs(r)
    assert c >= 0
    if a == 0:
        assert c == 0 and s == 1


# ---------------------------------------------------------------------------- plane wave (P2)

def _impedance_recursion(section, omega):
    """Surface impedance E/H of a layered earth by the textbook recursion (exp(-i omega t)):
    Z_n = omega mu0 / k_n, Z_j = zeta_j (Z_j+1 + zeta_j tanh(-i k_j h_j)) / (zeta_j + Z_j+1 tanh(-i k_j h_j)),
    k_j = sqrt(i omega mu0 zeta_j)."""
    zeta = section.zeta(omega)
    k = np.sqrt(1j * omega * cm.MU0 * zeta)
    zi = omega * cm.MU0 / k
    Z = zi[-1]
    for j in range(len(zeta) - 2, 0, -1):
        t = np.tanh(-1j * k[j] * section.thickness[j])
        Z = zi[j] * (Z + zi[j] * t) / (zi[j] + Z * t)
    return Z


def test_plane_wave_half_space(oracle):
    sigma = 0.01
    sec = cm.Section1D(0.0, [0.0, 1e5, 0.0], [0.0, sigma, sigma])   # a slab over its own half-space
    for f in (1e-3, 1.0, 100.0):
        w = cm.omega_of(f)
        Z = oracle.plane_wave_e(sec, w, 0.0) / oracle.plane_wave_h(sec, w, 0.0)
        rho_a = abs(Z) ** 2 / (w * cm.MU0)
        assert abs(rho_a * sigma - 1) < 1e-6, (f, rho_a)
        assert abs(np.angle(Z) + np.pi / 4) < 1e-6, (f, np.angle(Z))
        # E decays as exp(i k z) below the surface
        k = np.sqrt(1j * w * cm.MU0 * cm.zeta_of(sigma, w))
        z = 123.0
        ratio = oracle.plane_wave_e(sec, w, z) / oracle.plane_wave_e(sec, w, 0.0)
        assert abs(ratio - np.exp(1j * k * z)) < 1e-12


@pytest.mark.parametrize("thick", [[0.0, 500.0, 0.0], [0.0, 300.0, 2000.0, 0.0]])
def test_plane_wave_layered_impedance(oracle, thick):
    sig = [0.0, 0.1, 0.001] if len(thick) == 3 else [0.0, 0.02, 1.0, 0.003]
    sec = cm.Section1D(0.0, thick, sig)
    for f in (0.01, 1.0, 30.0):
        w = cm.omega_of(f)
        Z = oracle.plane_wave_e(sec, w, 0.0) / oracle.plane_wave_h(sec, w, 0.0)
        assert abs(Z / _impedance_recursion(sec, w) - 1) < 1e-12, f


def test_product_plane_wave_matches_oracle(oracle):
    """The product's host-side plane wave (used to build j_s) equals the oracle's."""
    from extreme_amd import forward as F
    m = cm.config_c2()
    for f in (1e-3, 1.0, 10.0):
        w = cm.omega_of(f)
        got = F.normal_field(m, w)
        want = oracle.normal_field(m, w)
        assert rel(got, want) < 1e-13
        for d in (0.0, 750.0, 3100.0):
            assert abs(F.plane_wave_h(m.section, w, d) / oracle.plane_wave_h(m.section, w, d) - 1) < 1e-13


# ---------------------------------------------------------------------------- Green tensor (G4-G12)

def _homogeneous(nz, dz=100.0, sigma=0.05, nx=4, ny=3):
    sec = cm.Section1D(0.0, [0.0, 5000.0, 0.0], [sigma, sigma, sigma])
    depth = 1000.0 + np.arange(nz) * dz
    thick = np.full(nz, dz)
    return cm.CartesianModel(nx, ny, 200.0, 150.0, depth, thick, np.full((nx, ny, nz), sigma), sec, "homog", [1.0])


def test_green_homogeneous_is_translation_invariant(oracle):
    """In a homogeneous medium with equal-thickness layers G(rc, tr) depends on rc - tr only, so the
    match/above/below branches of the U functions must agree; xz, yz are odd in rc - tr and
    xx, yy, zz, xy are even."""
    m = _homogeneous(4)
    G = oracle.green_atoa(m, cm.omega_of(2.0), fft=False)
    nz = m.nz
    L = 4 * m.nx * m.ny

    def full(c):
        v = G[c].reshape(L, -1)
        if c in ("xz", "yz"):
            return v.reshape(L, nz, nz)
        out = np.zeros((L, nz, nz), dtype=complex)
        for tr in range(nz):
            for rc in range(tr + 1):
                out[:, rc, tr] = out[:, tr, rc] = v[:, rc + tr * (tr + 1) // 2]
        return out
    for c in oracle.COMPONENTS:
        g = full(c)
        scale = np.max(np.abs(g))
        for d in range(-nz + 1, nz):
            diag = [g[:, r, r + d] if d >= 0 else g[:, r - d, r] for r in range(nz - abs(d))]
            for v in diag[1:]:
                assert np.max(np.abs(v - diag[0])) < 1e-9 * scale, (c, d)
        if c in ("xz", "yz"):
            assert np.max(np.abs(g + g.transpose(0, 2, 1))) < 1e-9 * scale, c
    # the tensor is not trivially zero and the vertical coupling is not trivially symmetric
    assert np.max(np.abs(full("xz"))) > 0


def test_green_mirror_parities(oracle):
    """The mirrored spatial tensor on the 2Nx x 2Ny grid (GreenTensorFftPopulator.cs:41-51): row
    Nx and column Ny are zero, and negative offsets carry the parities of each component."""
    m = cm.small_model(5, 4, 2, seed=3)
    G = oracle.green_atoa(m, cm.omega_of(1.0), fft=False)
    nx, ny = m.nx, m.ny
    for c, (px, py) in _PARITY.items():
        g = G[c].reshape(2 * nx, 2 * ny, -1)
        assert np.all(g[nx] == 0) and np.all(g[:, ny] == 0), c
        for i in range(1, nx):
            for j in range(0, ny):
                assert np.array_equal(g[2 * nx - i, j], px * g[i, j]), c
        for i in range(0, nx):
            for j in range(1, ny):
                assert np.array_equal(g[i, 2 * ny - j], py * g[i, j]), c
