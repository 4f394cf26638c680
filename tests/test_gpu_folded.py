"""Folded Green tensor storage: only the (Nx+1)(Ny+1) blocks kx <= Nx, ky <= Ny are kept on the
device, the other wavenumbers' blocks follow from the mirror symmetry of the lateral tensor.

The folded operator is checked against the FP64 oracle (which keeps all 4 Nx Ny blocks), against
the same library with folding switched off (XTR_GREEN_LAYOUT=full), and through the export path
(the unfolded tensor equals the oracle's). A tensor without the symmetry stays unfolded."""
import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import xtreme as X
from tests.conftest import rel

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def _op(oracle, m, monkeypatch=None, full=False, precision="fp64", G=None, max_rhs=2):
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w) if G is None else G
    if full:
        monkeypatch.setenv("XTR_GREEN_LAYOUT", "full")
    op = X.ConvolutionOperator(m, w, precision=precision, max_rhs=max_rhs)
    op.set_green_tensor(G)
    if full:
        monkeypatch.delenv("XTR_GREEN_LAYOUT")
    return w, G, op


@pytest.mark.parametrize("shape", [(4, 4, 3), (6, 4, 5), (8, 12, 2), (6, 4, 8), (5, 6, 12), (4, 6, 16), (6, 4, 24),
                                   (4, 4, 32), (6, 6, 32), (4, 4, 64)])
def test_folded_apply_matches_oracle_and_full(oracle, torch_cuda, monkeypatch, shape):
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nx * 7 + ny + nz)
    w, G, op = _op(oracle, m)
    assert op.green_layout() == ("folded", (nx + 1) * (ny + 1))
    _, _, opf = _op(oracle, m, monkeypatch, full=True, G=G)
    assert opf.green_layout() == ("full", 4 * nx * ny)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    oop = oracle.Operator(m, G)
    rng = np.random.default_rng(nz)
    xs = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    for kind in ("A", "Chi0"):
        op.prepare_operator(kind, z)
        opf.prepare_operator(kind, z)
        oop.prepare(kind, z, z0)
        ref = np.stack([oop.apply(xs[0]), oop.apply(xs[1])])
        y2, y1 = op.apply(xs), op.apply(xs[1])
        yf = opf.apply(xs)
        print(f"folded {shape} {kind}: vs oracle {rel(y2, ref):.2e} nrhs1 {rel(y1, ref[1]):.2e} "
              f"vs full {rel(y2, yf):.2e}", flush=True)
        assert rel(y2, ref) < TOL_FP64
        assert rel(y1, ref[1]) < TOL_FP64
        assert rel(y2, yf) < 1e-13


@pytest.mark.parametrize("shape,nrhs", [((6, 4, 8), 1), ((5, 6, 12), 2), ((4, 6, 16), 4), ((6, 4, 24), 8),
                                        ((5, 6, 5), 2), ((3, 5, 31), 2), ((5, 3, 1), 3)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_folded_small_nz_kernel_matches_generic(oracle, torch_cuda, monkeypatch, shape, nrhs, precision):
    """The folded contractions for nz <= 31 (one Green-block read for all images and right-hand sides:
    vector FMAs, or MFMA from nz = 17 to 24 by default and from nz = 4 with XTR_CONTRACT_KERNEL=mfma)
    against the generic per-wavenumber kernel and the oracle, 1-8 right-hand sides, complex128 and
    complex64."""
    torch = torch_cuda
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nz + 3 * nrhs)
    w, G, op = _op(oracle, m, precision=precision, max_rhs=nrhs)
    assert op.green_layout()[0] == "folded"
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    oop = oracle.Operator(m, G)
    rng = np.random.default_rng(nz + nrhs)
    xs = rng.standard_normal((nrhs, m.n)) + 1j * rng.standard_normal((nrhs, m.n))
    x32 = torch.from_numpy(xs).to("cuda").to(torch.complex64)
    for kind in ("A", "Chi0"):
        op.prepare_operator(kind, z)
        oop.prepare(kind, z, z0)
        ref = oop.apply(xs[-1])
        def run(kernel):
            if kernel:
                monkeypatch.setenv("XTR_CONTRACT_KERNEL", kernel)
            out = op.apply(xs) if precision == "fp64" else op.apply_c64(x32).cpu().numpy()
            if kernel:
                monkeypatch.delenv("XTR_CONTRACT_KERNEL")
            return out.reshape(nrhs, m.n)
        y, yv, ym, yg = run(None), run("vector"), run("mfma"), run("generic")
        tol_g, tol_o = (1e-13, 1e-12) if precision == "fp64" else (2e-6, 2e-6)
        print(f"small-nz folded {shape} nrhs {nrhs} {precision} {kind}: vs generic {rel(y, yg):.2e}, "
              f"vector {rel(yv, yg):.2e}, mfma {rel(ym, yg):.2e}, vs oracle {rel(y[-1], ref):.2e}", flush=True)
        assert rel(y, yg) < tol_g
        assert rel(yv, yg) < tol_g
        assert rel(ym, yg) < tol_g
        assert rel(y[-1], ref) < tol_o


@pytest.mark.parametrize("shape", [(4, 4, 3), (6, 4, 5)])
def test_folded_export_unfolds(oracle, torch_cuda, shape):
    """get_green of a folded tensor rebuilds all 4 Nx Ny blocks: bit-exact for power-of-two FFT
    lengths (the oracle's mirrored blocks are then exact), within FFT rounding otherwise."""
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=11)
    w, G, op = _op(oracle, m)
    back = op.green_tensor()
    pow2 = (nx & (nx - 1)) == 0 and (ny & (ny - 1)) == 0
    for c in oracle.COMPONENTS:
        b = back[c].reshape(G[c].shape)
        if pow2:
            assert np.array_equal(b, G[c]), c
        else:
            assert np.abs(b - G[c]).max() <= 1e-14 * np.abs(G[c]).max(), c


def test_random_tensor_stays_full(oracle, torch_cuda):
    m = cm.small_model(4, 4, 3, seed=2)
    w = cm.omega_of(m.frequencies[0])
    rng = np.random.default_rng(5)
    L, S = 4 * 16, 6
    G = {c: rng.standard_normal((L, 9 if c in ("xz", "yz") else S)) + 0j for c in oracle.COMPONENTS}
    op = X.ConvolutionOperator(m, w)
    op.set_green_tensor(G)
    assert op.green_layout() == ("full", L)


def test_device_built_green_is_folded(oracle, torch_cuda):
    """The device assembly stores the folded tensor; its export equals the oracle's tensor and the
    folded storage is a quarter of the full one (plus the kx = Nx and ky = Ny lines)."""
    m = cm.config_c1()
    w = cm.omega_of(m.frequencies[0])
    op = X.ConvolutionOperator(m, w)
    op.calculate_green_tensor()
    assert op.green_layout() == ("folded", (m.nx + 1) * (m.ny + 1))
    n, gbytes, _ = op.sizes()
    nz = m.nz
    assert gbytes == (m.nx + 1) * (m.ny + 1) * (4 * nz * (nz + 1) // 2 + 2 * nz * nz) * 16
    G = oracle.green_atoa(m, w)
    back = op.green_tensor()
    for c in oracle.COMPONENTS:
        e = np.abs(back[c].reshape(G[c].shape) - G[c]).max() / np.abs(G[c]).max()
        print(f"C1 device Green {c}: {e:.2e}", flush=True)
        assert e < 1e-12


def test_folded_fp32(oracle, torch_cuda):
    """complex64 copy of the folded tensor (mixed precision reads the same copy)."""
    m = cm.small_model(6, 6, 32, seed=4)
    w, G, op = _op(oracle, m, precision="fp32")
    assert op.green_layout()[0] == "folded"
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    oop = oracle.Operator(m, G)
    op.prepare_operator("A", z)
    oop.prepare("A", z, z0)
    x = np.random.default_rng(1).standard_normal((2, m.n)) + 0.25j
    ref = np.stack([oop.apply(x[0]), oop.apply(x[1])])
    e = rel(op.apply(x), ref)
    print(f"folded fp32: {e:.2e}", flush=True)
    assert e < 2e-6


def _mirror_symmetric_green(oracle, model, seed):
    """A random Green tensor with the mirror symmetry of the lateral tensor: blocks kx <= Nx,
    ky <= Ny drawn at random, the others their mirror images (xy negated under either mirror alone,
    xz under kx, yz under ky), in the reference layout."""
    nx, ny, nz = model.nx, model.ny, model.nz
    rng = np.random.default_rng(seed)
    S, A = nz * (nz + 1) // 2, nz * nz
    sizes = {"xx": S, "xy": S, "yy": S, "zz": S, "xz": A, "yz": A}
    src = {c: rng.standard_normal(((nx + 1) * (ny + 1), p)) + 1j * rng.standard_normal(((nx + 1) * (ny + 1), p))
           for c, p in sizes.items()}
    kx, ky = np.divmod(np.arange(4 * nx * ny), 2 * ny)
    mx, my = kx > nx, ky > ny
    b = np.where(mx, 2 * nx - kx, kx) * (ny + 1) + np.where(my, 2 * ny - ky, ky)
    sign = {"xx": 1, "yy": 1, "zz": 1, "xy": np.where(mx ^ my, -1, 1), "xz": np.where(mx, -1, 1),
            "yz": np.where(my, -1, 1)}
    return {c: src[c][b] * np.reshape(sign[c], (-1, 1)) for c in oracle.COMPONENTS}


@pytest.mark.parametrize("shape", [(24, 24, 32), (16, 16, 64)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_folded_several_blocks_per_workgroup(oracle, torch_cuda, shape, precision):
    """More folded blocks than workgroups (the copy of a workgroup's next block overlaps the current
    one), one, two and three right-hand sides (three: a second launch with one)."""
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nz + 1)
    w = cm.omega_of(m.frequencies[0])
    G = _mirror_symmetric_green(oracle, m, seed=nx + nz)
    op = X.ConvolutionOperator(m, w, precision=precision, max_rhs=3)
    op.set_green_tensor(G)
    assert op.green_layout() == ("folded", (nx + 1) * (ny + 1))
    oop = oracle.Operator(m, G)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op.prepare_operator("A", z)
    oop.prepare("A", z, z0)
    rng = np.random.default_rng(29)
    xs = rng.standard_normal((3, m.n)) + 1j * rng.standard_normal((3, m.n))
    ref = np.stack([oop.apply(x) for x in xs])
    tol = TOL_FP64 if precision == "fp64" else 2e-6
    errs = [rel(op.apply(xs[:k]), ref[:k]) for k in (1, 2, 3)]
    print(f"folded {shape} {precision}: nrhs 1/2/3 {errs[0]:.2e} {errs[1]:.2e} {errs[2]:.2e}", flush=True)
    assert max(errs) < tol
