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


def _op(oracle, m, monkeypatch=None, full=False, precision="fp64", G=None):
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w) if G is None else G
    if full:
        monkeypatch.setenv("XTR_GREEN_LAYOUT", "full")
    op = X.ConvolutionOperator(m, w, precision=precision, max_rhs=2)
    op.set_green_tensor(G)
    if full:
        monkeypatch.delenv("XTR_GREEN_LAYOUT")
    return w, G, op


@pytest.mark.parametrize("shape", [(4, 4, 3), (6, 4, 5), (8, 12, 2), (4, 4, 32), (6, 6, 32), (4, 4, 64)])
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
