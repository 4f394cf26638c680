"""GPU parity of the operator apply (A9) against the FP64 oracle, on oracle-built Green tensors."""
import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import xtreme as X
from tests.conftest import rel

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12   # FP64 apply vs FP64 oracle, relative 2-norm
TOL_FP32 = 2e-6    # complex64 Green tensor and FFT


def _setup(oracle, model, precision="fp64", max_rhs=2):
    w = cm.omega_of(model.frequencies[0])
    G = oracle.green_atoa(model, w)
    op = X.ConvolutionOperator(model, w, precision=precision, max_rhs=max_rhs)
    op.set_green_tensor(G)
    oop = oracle.Operator(model, G)
    return w, G, op, oop


@pytest.mark.parametrize("shape", [(4, 4, 2), (8, 12, 3), (16, 8, 5), (4, 4, 24), (4, 4, 33), (4, 4, 64),
                                   (5, 4, 3), (25, 15, 2), (40, 40, 2)])
@pytest.mark.parametrize("kind", ["A", "Chi0"])
def test_apply_small_grids(oracle, torch_cuda, shape, kind):
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nx * 100 + ny * 10 + nz)
    w, G, op, oop = _setup(oracle, m)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op.prepare_operator(kind, z)
    oop.prepare(kind, z, z0)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    y = op.apply(x)
    print(f"apply {kind} {shape}: rel {rel(y, oop.apply(x)):.3e}", flush=True)
    assert rel(y, oop.apply(x)) < TOL_FP64


def test_apply_c1_host_and_device(oracle, torch_cuda):
    torch = torch_cuda
    m = cm.config_c1()
    w, G, op, oop = _setup(oracle, m)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    rng = np.random.default_rng(3)
    xs = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    for kind in ("A", "Chi0"):
        op.prepare_operator(kind, z)
        oop.prepare(kind, z, z0)
        ref = np.stack([oop.apply(xs[0]), oop.apply(xs[1])])
        yh = op.apply(xs)                                         # host, nrhs = 2
        xd = torch.from_numpy(xs).to("cuda")
        yd = op.apply(xd).cpu().numpy()                           # device, nrhs = 2
        y1 = op.apply(xs[1])                                      # nrhs = 1
        print(f"C1 {kind}: host {rel(yh, ref):.3e} device {rel(yd, ref):.3e}", flush=True)
        assert rel(yh, ref) < TOL_FP64
        assert rel(yd, ref) < TOL_FP64
        assert rel(y1, ref[1]) < TOL_FP64
    # a result that overlaps an input (in place, or shifted by half a vector) is refused before
    # any launch, and the input stays intact
    buf = torch.zeros(3 * m.n, dtype=torch.complex128, device="cuda")
    buf[:2 * m.n] = torch.from_numpy(xs.reshape(-1)).to("cuda")
    x2 = buf[:2 * m.n].view(2, m.n)
    for x, y in ((x2, x2), (x2[0], buf[m.n // 2:m.n // 2 + m.n]), (x2, buf[m.n:3 * m.n].view(2, m.n))):
        with pytest.raises(X.XtrError, match="overlaps"):
            op.apply(x, y)
    assert np.array_equal(x2.cpu().numpy(), xs)


def test_identity_when_zeta_equals_background(oracle, torch_cuda):
    """zeta == zeta0 makes R = 0, so A = I exactly and Chi0 = 0 (ConvolutionOperator.cs:138-141)."""
    m = cm.config_c1()
    w, G, op, oop = _setup(oracle, m)
    zb = np.repeat(m.zeta0(w)[None, :], m.nx * m.ny, axis=0).reshape(-1)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    op.prepare_operator("A", zb)
    assert np.array_equal(op.apply(x), x)


def test_apply_fp32_error_bound(oracle, torch_cuda):
    m = cm.config_c1()
    w, G, op, oop = _setup(oracle, m, precision="fp32")
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op.prepare_operator("A", z)
    oop.prepare("A", z, z0)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    e = rel(op.apply(x), oop.apply(x))
    print(f"C1 fp32 apply rel error {e:.3e}", flush=True)
    assert e < TOL_FP32


def _random_green(oracle, model, seed):
    """A seeded random Green tensor in the reference layout (exercises every element position)."""
    rng = np.random.default_rng(seed)
    L = 4 * model.nx * model.ny
    nz = model.nz
    S = nz * (nz + 1) // 2
    return {c: rng.standard_normal((L, nz * nz if c in ("xz", "yz") else S))
            + 1j * rng.standard_normal((L, nz * nz if c in ("xz", "yz") else S)) for c in oracle.COMPONENTS}


@pytest.mark.parametrize("shape", [(4, 4, 32), (12, 6, 32), (24, 12, 32)])
def test_apply_nz32_random_green(oracle, torch_cuda, shape):
    """nz = 32 runs the streaming contraction: fewer, equal and more wavenumbers than CUs, an uneven
    split over the persistent workgroups, one and two right-hand sides."""
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nx + ny)
    w = cm.omega_of(m.frequencies[0])
    G = _random_green(oracle, m, seed=nx * ny)
    op = X.ConvolutionOperator(m, w, max_rhs=2)
    op.set_green_tensor(G)
    oop = oracle.Operator(m, G)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op.prepare_operator("A", z)
    oop.prepare("A", z, z0)
    rng = np.random.default_rng(11)
    xs = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    ref = np.stack([oop.apply(xs[0]), oop.apply(xs[1])])
    y2 = op.apply(xs)
    y1 = op.apply(xs[0])
    print(f"nz32 {shape}: nrhs2 {rel(y2, ref):.3e} nrhs1 {rel(y1, ref[0]):.3e}", flush=True)
    assert rel(y2, ref) < TOL_FP64
    assert rel(y1, ref[0]) < TOL_FP64


def test_streaming_contraction_matches_generic(oracle, torch_cuda, monkeypatch):
    """The two contraction kernels agree on the same operator and input (summation order only)."""
    m = cm.small_model(24, 12, 32, seed=5)
    w = cm.omega_of(m.frequencies[0])
    G = _random_green(oracle, m, seed=3)
    op = X.ConvolutionOperator(m, w, max_rhs=2)
    op.set_green_tensor(G)
    op.prepare_operator("A")
    xs = np.random.default_rng(2).standard_normal((2, m.n)) + 0.5j
    y_stream = op.apply(xs)
    monkeypatch.setenv("XTR_CONTRACT_KERNEL", "generic")
    y_generic = op.apply(xs)
    monkeypatch.delenv("XTR_CONTRACT_KERNEL")
    print(f"stream vs generic: {rel(y_stream, y_generic):.3e}", flush=True)
    assert rel(y_stream, y_generic) < 1e-14


@pytest.mark.parametrize("shape", [(4, 4, 64), (12, 12, 64)])
def test_apply_nz64_streaming(oracle, torch_cuda, monkeypatch, shape):
    """nz = 64 runs the four-unit streaming contraction (the Green block does not fit LDS): fewer
    wavenumbers than CUs and an uneven split, one and two right-hand sides, against the oracle and
    against the generic kernel."""
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nx * 3 + ny)
    w = cm.omega_of(m.frequencies[0])
    G = _random_green(oracle, m, seed=nx * ny + 64)
    op = X.ConvolutionOperator(m, w, max_rhs=2)
    op.set_green_tensor(G)
    oop = oracle.Operator(m, G)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op.prepare_operator("A", z)
    oop.prepare("A", z, z0)
    rng = np.random.default_rng(17)
    xs = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    ref = np.stack([oop.apply(xs[0]), oop.apply(xs[1])])
    y2 = op.apply(xs)
    y1 = op.apply(xs[1])
    monkeypatch.setenv("XTR_CONTRACT_KERNEL", "generic")
    yg = op.apply(xs)
    monkeypatch.delenv("XTR_CONTRACT_KERNEL")
    print(f"nz64 {shape}: nrhs2 {rel(y2, ref):.3e} nrhs1 {rel(y1, ref[1]):.3e} "
          f"vs generic {rel(y2, yg):.3e}", flush=True)
    assert rel(y2, ref) < TOL_FP64
    assert rel(y1, ref[1]) < TOL_FP64
    assert rel(y2, yg) < 1e-14


@pytest.mark.parametrize("nz", [5, 8, 24, 32, 64])
def test_green_set_get_roundtrip(oracle, torch_cuda, nz):
    """Reference layout -> device layout (swizzled asymmetric blocks) -> reference layout, bit-exact."""
    m = cm.small_model(3, 2, nz, seed=nz)
    w = cm.omega_of(m.frequencies[0])
    G = _random_green(oracle, m, seed=nz)
    op = X.ConvolutionOperator(m, w)
    op.set_green_tensor(G)
    back = op.green_tensor()
    for c in oracle.COMPONENTS:
        assert np.array_equal(back[c].reshape(G[c].shape), G[c]), c


@pytest.mark.parametrize("shape", [(4, 4, 32), (24, 12, 32), (4, 4, 64), (12, 12, 64)])
def test_apply_fp32_streaming(oracle, torch_cuda, monkeypatch, shape):
    """complex64 streaming contractions (nz = 32: whole blocks through a 4-buffer ring; nz = 64: four
    units per block through a 4-slot ring): within the FP32 bound of the FP64 oracle and equal to the
    complex64 generic kernel up to summation order."""
    nx, ny, nz = shape
    m = cm.small_model(nx, ny, nz, seed=nx * ny)
    w = cm.omega_of(m.frequencies[0])
    G = _random_green(oracle, m, seed=nx + 7)
    op = X.ConvolutionOperator(m, w, precision="fp32", max_rhs=2)
    op.set_green_tensor(G)
    oop = oracle.Operator(m, G)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    op.prepare_operator("A", z)
    oop.prepare("A", z, z0)
    xs = np.random.default_rng(8).standard_normal((2, m.n)) + 1j * np.random.default_rng(9).standard_normal((2, m.n))
    ref = np.stack([oop.apply(xs[0]), oop.apply(xs[1])])
    y2 = op.apply(xs)
    y1 = op.apply(xs[1])
    monkeypatch.setenv("XTR_CONTRACT_KERNEL", "generic")
    yg = op.apply(xs)
    monkeypatch.delenv("XTR_CONTRACT_KERNEL")
    print(f"fp32 {shape}: vs oracle {rel(y2, ref):.2e} nrhs1 {rel(y1, ref[1]):.2e} "
          f"vs generic {rel(y2, yg):.2e}", flush=True)
    assert rel(y2, ref) < TOL_FP32
    assert rel(y1, ref[1]) < TOL_FP32
    assert rel(y2, yg) < 1e-6


@pytest.mark.parametrize("tiles", ["nnnn", "wwww", "mmmm", "rrrr", "ssss", None],
                         ids=["narrow", "wide", "narrow512", "reg", "reg_alt", "tuned"])
def test_fft_tiles_give_the_same_apply(oracle, torch_cuda, monkeypatch, tiles):
    """Every FFT column tile (forced, or the measured choice) gives the same apply as the oracle.
    Tiles n, w, m only regroup columns, so FP64 results stay within rounding of the default tile;
    the register tiles r, s also change the radix plan (16 x 8 instead of 4 x 4 x 4 x 2 for 128
    points), which moves results by a few units in the last place."""
    m = cm.small_model(64, 64, 2, seed=21)                 # pencils of 128: all four tiles exist
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    oop = oracle.Operator(m, G)
    oop.prepare("A", z, z0)
    rng = np.random.default_rng(3)
    xs = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    ref = np.stack([oop.apply(x) for x in xs])
    monkeypatch.setenv("XTR_FFT_TILES", "dddd")
    op = X.ConvolutionOperator(m, w, max_rhs=2)
    op.set_green_tensor(G)
    op.prepare_operator("A", z)
    y_default = op.apply(xs)
    if tiles is None:
        monkeypatch.delenv("XTR_FFT_TILES")
    else:
        monkeypatch.setenv("XTR_FFT_TILES", tiles)
    for rep in range(3):                                    # the first apply of a shape measures
        y = op.apply(xs)
        e, d = rel(y, ref), rel(y, y_default)
        print(f"tiles {tiles} rep {rep}: {op.fft_tiles()} vs oracle {e:.2e}, vs default tile {d:.2e}", flush=True)
        assert e < 1e-12 and d < (1e-14 if tiles and tiles[0] in "nwm" else 1e-13)
    got = op.fft_tiles()
    assert set(got) == {"fwd_y", "fwd_x", "inv_x", "inv_y"}
    if tiles is not None:
        want = {"n": "narrow", "w": "wide", "m": "narrow512", "r": "reg", "s": "reg_alt"}[tiles[0]]
        assert all(v == want for v in got.values())


@pytest.mark.parametrize("shape", [(32, 128, 3), (256, 16, 2)], ids=["p64_p256", "p512"])
@pytest.mark.parametrize("tiles", ["rrrr", "ssss"], ids=["reg", "reg_alt"])
def test_register_fft_tiles_other_lengths(oracle, torch_cuda, monkeypatch, shape, tiles):
    """The register pencils at 64 and 256 points (radix 8 x 8, 16 x 16) and at 512 (8 x 8 x 8, two
    LDS exchanges), FP64 and FP32 applies against the oracle."""
    m = cm.small_model(*shape, seed=5)
    w = cm.omega_of(m.frequencies[0])
    G = oracle.green_atoa(m, w)
    z, z0 = m.zeta(w).reshape(-1), m.zeta0(w)
    oop = oracle.Operator(m, G)
    oop.prepare("A", z, z0)
    rng = np.random.default_rng(4)
    xs = rng.standard_normal((2, m.n)) + 1j * rng.standard_normal((2, m.n))
    ref = np.stack([oop.apply(x) for x in xs])
    monkeypatch.setenv("XTR_FFT_TILES", tiles)
    for prec, tol in (("fp64", 1e-12), ("fp32", TOL_FP32)):
        op = X.ConvolutionOperator(m, w, max_rhs=2, precision=prec)
        op.set_green_tensor(G)
        op.prepare_operator("A", z)
        e = rel(op.apply(xs), ref)
        print(f"{shape} {tiles} {prec}: {op.fft_tiles()} vs oracle {e:.2e}", flush=True)
        assert e < tol


def test_hbm_stream_probe_writes_its_blocks(torch_cuda):
    """The stream probe (a measurement aid) reads every block and writes write_kb KB after each: the
    second double of element e of block b's output is e, and nothing past blocks * write_kb KB is
    touched."""
    torch = torch_cuda
    blocks, read_kb, write_kb = 300, 64, 2
    src = torch.ones(blocks * read_kb * 128, dtype=torch.float64, device="cuda")
    dst = torch.full((blocks * write_kb * 128 + 256,), -7.0, dtype=torch.float64, device="cuda")
    X.probe_stream(src, dst, blocks, read_kb, write_kb)
    torch.cuda.synchronize()
    out = dst[: blocks * write_kb * 128].view(blocks, write_kb * 64, 2)
    e = torch.arange(write_kb * 64, dtype=torch.float64, device="cuda")
    assert torch.equal(out[:, :, 1], e.expand(blocks, -1))
    assert torch.all(dst[blocks * write_kb * 128:] == -7.0)
    with pytest.raises(X.XtrError):
        X.probe_stream(src, dst, blocks, 48, write_kb)               # not a multiple of 32 KB
    with pytest.raises(ValueError):
        X.probe_stream(src, dst, blocks + 1, read_kb, write_kb)      # src too small
