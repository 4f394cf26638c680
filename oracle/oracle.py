"""ctypes binding of the FP64 CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Only the tests, the smoke
check and the CPU-baseline leg of bench.py import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class CDouble(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Section(C.Structure):
    _fields_ = [("nlayers", C.c_int), ("zero_level", C.c_double),
                ("thickness", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double))]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("depth", C.POINTER(C.c_double)), ("thick", C.POINTER(C.c_double))]


class GmresParams(C.Structure):
    _fields_ = [("n", C.c_int64), ("m_outer", C.c_int), ("m_inner", C.c_int), ("max_repeat", C.c_int),
                ("tol", C.c_double), ("backward_check", C.c_int)]


class Result(C.Structure):
    _fields_ = [("converged", C.c_int), ("bea", C.c_double), ("be", C.c_double), ("iterations", C.c_int),
                ("inner_iterations", C.c_int), ("matvecs", C.c_int64), ("dots", C.c_int64)]


class Op(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + \
               [(n, C.c_void_p) for n in ("xx", "xy", "xz", "yy", "yz", "zz", "r", "f", "b")] + \
               [("type", C.c_int), ("buf1", C.c_void_p), ("buf2", C.c_void_p), ("rx", C.c_void_p)]


MATVEC = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p)
ITER = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.xo_zeta.restype = CDouble
        L.xo_zeta.argtypes = [C.c_double, C.c_double]
        L.xo_plane_wave_e.restype = CDouble
        L.xo_plane_wave_e.argtypes = [C.POINTER(Section), C.c_double, C.c_double, C.c_double]
        L.xo_plane_wave_h.restype = CDouble
        L.xo_plane_wave_h.argtypes = [C.POINTER(Section), C.c_double, C.c_double]
        L.xo_fft2_batch.argtypes = [c128p, C.c_int, C.c_int, C.c_int64, C.c_int]
        L.xo_fft1.argtypes = [c128p, C.c_int, C.c_int]
        L.xo_green_atoa.argtypes = [C.POINTER(Section), C.POINTER(Grid), C.c_double, C.c_int, f64p, f64p,
                                    C.c_int] + [c128p] * 6
        L.xo_green_scalars.argtypes = [C.POINTER(Section), C.POINTER(Grid), C.c_double, f64p, f64p,
                                       C.c_void_p, C.c_int, c128p]
        L.xo_plan_create.argtypes = [C.POINTER(Grid), C.c_int, C.c_void_p]
        L.xo_plan_free.argtypes = [C.c_void_p]
        L.xo_contract.argtypes = [C.c_int, C.c_int64] + [c128p] * 8
        L.xo_op_init.argtypes = [C.POINTER(Op), C.c_int, C.c_int, C.c_int] + [c128p] * 6
        L.xo_op_prepare.argtypes = [C.POINTER(Op), C.c_int, c128p, c128p, f64p]
        L.xo_op_apply.argtypes = [C.POINTER(Op), C.c_void_p, C.c_void_p]
        L.xo_op_free.argtypes = [C.POINTER(Op)]
        L.xo_fgmres_solve.argtypes = [C.POINTER(GmresParams), MATVEC, C.c_void_p, ITER, C.c_void_p,
                                      c128p, c128p, C.POINTER(Result)]
        L.xo_gmres_inner.argtypes = [C.POINTER(GmresParams), MATVEC, C.c_void_p, c128p, c128p,
                                     C.POINTER(Result)]
        L.xo_zrotg.argtypes = [CDouble, CDouble, C.POINTER(C.c_double), C.POINTER(CDouble)]
        L.xo_set_threads.argtypes = [C.c_int]
        L.xo_mu0.restype = C.c_double
        L.xo_green_atoo_level.argtypes = [C.POINTER(Section), C.POINTER(Grid), C.c_double, C.c_int, f64p, f64p,
                                          C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.POINTER(C.c_void_p)]
        L.xo_green_atoo_level_plan.argtypes = [C.POINTER(Section), C.POINTER(Grid), C.c_double, C.c_int, f64p,
                                               f64p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                               C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.xo_atoo_level_rho_range.argtypes = [C.POINTER(Grid), C.c_double, C.c_double, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]
        L.xo_atoo_level_scalars.argtypes = [C.POINTER(Section), C.POINTER(Grid), C.c_double, C.c_int, f64p, f64p,
                                            C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, c128p, C.c_int,
                                            f64p]
        L.xo_plan_create_range.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p]
        L.xo_atoo_field.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.c_int, c128p, c128p]
        L.xo_background_layer.restype = C.c_int
        L.xo_background_layer.argtypes = [C.POINTER(Section), C.c_double, C.c_double]
        L.xo_layer_at_depth.restype = C.c_int
        L.xo_layer_at_depth.argtypes = [C.POINTER(Section), C.c_double]
        L.xo_log10_rho_step.restype = C.c_double
        L.xo_hankel_lambdas.restype = C.c_int
        L.xo_hankel_lambdas.argtypes = [C.c_double, C.c_int, f64p]
        L.xo_hankel_filter.argtypes = [f64p, c128p, C.c_int, c128p]
        _lib = L
    return _lib


def set_threads(n: int):
    lib().xo_set_threads(int(n))


def _cplx(v):
    return complex(v.re, v.im)


class _Keep:
    """Holds numpy arrays alive for ctypes structures that point into them."""

    def __init__(self):
        self.refs = []

    def d(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.refs.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))


def make_section(sec, keep):
    return Section(sec.nlayers, float(sec.zero_level), keep.d(sec.thickness), keep.d(sec.sigma))


def make_grid(model, keep):
    return Grid(model.nx, model.ny, model.nz, float(model.dx), float(model.dy),
                keep.d(model.depth), keep.d(model.thick))


def zeta(sigma, omega):
    """zeta = sigma - i omega eps0 of one layer (OmegaModelUtils.cs:41-42)."""
    return _cplx(lib().xo_zeta(float(sigma), float(omega)))


def background_layer(section, start, end):
    """ModelUtils.FindCorrespondingBackgroundLayerIndex (Model/ModelUtils.cs:35-52); -1 when no finite
    layer holds [start, end] (the reference throws)."""
    k = _Keep()
    return lib().xo_background_layer(C.byref(make_section(section, k)), float(start), float(end))


def layer_at_depth(section, depth):
    """ScalarMathUtils.DetermineCorrespondingBackgroundLayer, half-open (ScalarMathUtils.cs:15-31); -1
    when no layer holds the depth (the reference throws)."""
    k = _Keep()
    return lib().xo_layer_at_depth(C.byref(make_section(section, k)), float(depth))


def plane_wave_e(section, omega, depth, thick=0.0):
    k = _Keep()
    s = make_section(section, k)
    v = _cplx(lib().xo_plane_wave_e(C.byref(s), omega, depth, thick))
    if np.isnan(v.real):
        raise ValueError("depth is not within the 1-D section")
    return v


def plane_wave_h(section, omega, depth):
    k = _Keep()
    s = make_section(section, k)
    v = _cplx(lib().xo_plane_wave_h(C.byref(s), omega, depth))
    if np.isnan(v.real):
        raise ValueError("depth is not within the 1-D section")
    return v


def fft2_batch(buf, n1, n2, nc, sign):
    out = np.ascontiguousarray(buf, dtype=np.complex128).copy()
    lib().xo_fft2_batch(out.reshape(-1), n1, n2, nc, sign)
    return out


def fft1(x, sign=-1):
    out = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().xo_fft1(out, len(out), sign)
    return out


def hankel_table():
    """The ndec=40 J0/J1 filter weights (data copied from the reference's hankelcoeff.xml)."""
    path = os.path.join(os.path.dirname(_HERE), "extreme_amd", "data", "hankel_ndec40.txt")
    t = np.loadtxt(path)
    return np.ascontiguousarray(t[:, 0]), np.ascontiguousarray(t[:, 1])


def hankel_rho_step():
    """(float)10^(1/40) as the reference rounds it (HankelCoefficients.cs:67)."""
    return lib().xo_log10_rho_step()


def hankel_lambdas(rho0, nrho):
    """lambda grid of the output grid rho[i] = rho0 * step^i (ScalarPlan.cs:76-97)."""
    lam = np.zeros(431 + nrho)
    n = lib().xo_hankel_lambdas(rho0, nrho, lam)
    return lam[:n]


def hankel_filter(hank, fk, nrho):
    """out[i] = sum_j hank[430 - j] fk[i + j] (HankelCoefficients.ConvoluteWithFast :103-118)."""
    fk = np.ascontiguousarray(fk, dtype=np.complex128)
    out = np.zeros(nrho, dtype=np.complex128)
    lib().xo_hankel_filter(np.ascontiguousarray(hank, dtype=np.float64), fk, nrho, out)
    return out


class _Plan(C.Structure):
    _fields_ = [("nitems", C.c_int), ("nrho", C.POINTER(C.c_int)), ("rho", C.POINTER(C.POINTER(C.c_double))),
                ("nlam", C.POINTER(C.c_int)), ("lam", C.POINTER(C.POINTER(C.c_double))),
                ("nrho_merged", C.c_int), ("rho_merged", C.POINTER(C.c_double)),
                ("merge_item", C.POINTER(C.c_int)), ("merge_index", C.POINTER(C.c_int)), ("zero_rho", C.c_int)]


def scalar_plan(model, number_of_hankels=10, rho_range=None):
    """ScalarPlansCreater.CreateForAnomalyToAnomaly, or CreateFor with rho_range = (rho_min,
    rho_max, zero_rho): returns the items' rho and lambda and the merged rho (rho = 0 first, item
    -1, with zero_rho)."""
    k = _Keep()
    g = make_grid(model, k)
    p = _Plan()
    if rho_range is None:
        rc = lib().xo_plan_create(C.byref(g), number_of_hankels, C.byref(p))
    else:
        rc = lib().xo_plan_create_range(number_of_hankels, float(rho_range[0]), float(rho_range[1]),
                                        1 if rho_range[2] else 0, C.byref(p))
    if rc:
        raise RuntimeError(f"xo_plan_create failed: {rc}")
    rho = [np.ctypeslib.as_array(p.rho[i], (p.nrho[i],)).copy() for i in range(p.nitems)]
    lam = [np.ctypeslib.as_array(p.lam[i], (p.nlam[i],)).copy() for i in range(p.nitems)]
    merged = np.ctypeslib.as_array(p.rho_merged, (p.nrho_merged,)).copy()
    item = np.ctypeslib.as_array(p.merge_item, (p.nrho_merged,)).copy()
    index = np.ctypeslib.as_array(p.merge_index, (p.nrho_merged,)).copy()
    lib().xo_plan_free(C.byref(p))
    return {"rho": rho, "lam": lam, "rho_merged": merged, "merge_item": item, "merge_index": index}


def green_scalars(model, omega, number_of_hankels=10, asym=True):
    """Green scalars on the merged rho grid: asym -> I3 [pairs][rho]; symm -> I1, I2, I5 [3][pairs][rho]."""
    h0, h1 = hankel_table()
    k = _Keep()
    s = make_section(model.section, k)
    g = make_grid(model, k)
    p = _Plan()
    rc = lib().xo_plan_create(C.byref(g), number_of_hankels, C.byref(p))
    if rc:
        raise RuntimeError(f"xo_plan_create failed: {rc}")
    nz = model.nz
    npairs = nz * nz if asym else nz * (nz + 1) // 2
    ncomp = 1 if asym else 3
    out = np.zeros(ncomp * npairs * p.nrho_merged, dtype=np.complex128)
    rc = lib().xo_green_scalars(C.byref(s), C.byref(g), omega, h0, h1, C.addressof(p), 1 if asym else 0, out)
    nr = p.nrho_merged
    lib().xo_plan_free(C.byref(p))
    if rc:
        raise RuntimeError("xo_green_scalars failed")
    return out.reshape(ncomp, npairs, nr)


COMPONENTS = ("xx", "xy", "xz", "yy", "yz", "zz")


def green_atoa(model, omega, number_of_hankels=10, fft=True):
    """AtoA Green tensor, reference layout: dict comp -> [2Nx*2Ny, P] complex128."""
    h0, h1 = hankel_table()
    k = _Keep()
    s = make_section(model.section, k)
    g = make_grid(model, k)
    nz = model.nz
    L = 4 * model.nx * model.ny
    S = nz * (nz + 1) // 2
    out = {c: np.zeros(L * (nz * nz if c in ("xz", "yz") else S), dtype=np.complex128) for c in COMPONENTS}
    rc = lib().xo_green_atoa(C.byref(s), C.byref(g), omega, number_of_hankels, h0, h1, 1 if fft else 0,
                             *[out[c] for c in COMPONENTS])
    if rc:
        raise RuntimeError(f"xo_green_atoa failed: {rc}")
    return {c: v.reshape(L, -1) for c, v in out.items()}


AO_COMPONENTS = ("xx", "yy", "zz", "xy", "yx", "xz", "yz", "zx", "zy")
AO_ELECTRIC = ("xx", "yy", "zz", "xy", "xz", "yz", "zx", "zy")
AO_MAGNETIC = ("xx", "xy", "yx", "xz", "yz", "zx", "zy")


def atoo_level_rho_range(model, x_shift=0.0, y_shift=0.0):
    """rho range of the level's scalar plan (ScalarPlansCreater.CreateForAnomalyToObservationLevels,
    :51-66): (rho_min, rho_max, zero_rho)."""
    k = _Keep()
    g = make_grid(model, k)
    rmin, rmax = C.c_double(), C.c_double()
    zero = lib().xo_atoo_level_rho_range(C.byref(g), float(x_shift), float(y_shift), C.byref(rmin), C.byref(rmax))
    return rmin.value, rmax.value, bool(zero)


def atoo_level_scalars(model, omega, level_z, rho_range, electric=True, number_of_hankels=10):
    """The level's scalars (PlanCalculator.CalculateForE/H with GreenScalarCalculator.CalculateZeroRho)
    on the merged rho grid of the plan rho_range = (rho_min, rho_max, zero_rho): (rho [n],
    I [ncomp, nz, n])."""
    h0, h1 = hankel_table()
    k = _Keep()
    s = make_section(model.section, k)
    g = make_grid(model, k)
    ncomp = 5 if electric else 4
    cap = 8192
    out = np.zeros(ncomp * model.nz * cap, dtype=np.complex128)
    rho = np.zeros(cap)
    n = lib().xo_atoo_level_scalars(C.byref(s), C.byref(g), omega, number_of_hankels, h0, h1, float(level_z),
                                    float(rho_range[0]), float(rho_range[1]), 1 if rho_range[2] else 0,
                                    1 if electric else 0, out, cap, rho)
    if n <= 0:
        raise RuntimeError(f"xo_atoo_level_scalars failed: {n}")
    return rho[:n].copy(), out[:ncomp * model.nz * n].reshape(ncomp, model.nz, n)


def green_atoo_level(model, omega, level_z=0.0, electric=True, number_of_hankels=10, fft=True, x_shift=0.0,
                     y_shift=0.0, rho_range=None):
    """AtoO level tensor (AtoOCalculator.CalculateGreenTensorAtoOLevel{Electric,Magnetic}) of the
    level at depth level_z whose sites are the cell centres shifted by (x_shift, y_shift) metres:
    dict comp -> [2Nx*2Ny, Nz] for the components the reference computes. rho_range =
    (rho_min, rho_max, zero_rho) replaces the level's own scalar plan range."""
    h0, h1 = hankel_table()
    k = _Keep()
    s = make_section(model.section, k)
    g = make_grid(model, k)
    L = 4 * model.nx * model.ny
    comps = AO_ELECTRIC if electric else AO_MAGNETIC
    out = {c: np.zeros(L * model.nz, dtype=np.complex128) for c in comps}
    ptrs = (C.c_void_p * 9)(*[out[c].ctypes.data if c in out else None for c in AO_COMPONENTS])
    if rho_range is None:
        rc = lib().xo_green_atoo_level(C.byref(s), C.byref(g), omega, number_of_hankels, h0, h1, float(level_z),
                                       float(x_shift), float(y_shift), 1 if electric else 0, 1 if fft else 0, ptrs)
    else:
        rc = lib().xo_green_atoo_level_plan(C.byref(s), C.byref(g), omega, number_of_hankels, h0, h1,
                                            float(level_z), float(x_shift), float(y_shift), float(rho_range[0]),
                                            float(rho_range[1]), 1 if rho_range[2] else 0, 1 if electric else 0,
                                            1 if fft else 0, ptrs)
    if rc:
        raise RuntimeError(f"xo_green_atoo_level failed: {rc}")
    return {c: v.reshape(L, model.nz) for c, v in out.items()}


def atoo_field(model, green, electric, jq):
    """Anomalous E (electric) or H field at the level, [Nx, Ny, 3] (ToOCalculator.cs:21-45)."""
    keep = {c: np.ascontiguousarray(green[c], dtype=np.complex128).reshape(-1) for c in green}
    ptrs = (C.c_void_p * 9)(*[keep[c].ctypes.data if c in keep else None for c in AO_COMPONENTS])
    jq = np.ascontiguousarray(jq, dtype=np.complex128).reshape(-1)
    out = np.zeros(model.nx * model.ny * 3, dtype=np.complex128)
    lib().xo_atoo_field(model.nx, model.ny, model.nz, ptrs, 1 if electric else 0, jq, out)
    return out.reshape(model.nx, model.ny, 3)


def impedance(e1, e2, h1, h2):
    """ResponseFunctionsCalculator.CalculateTensor (ResponseFunctionsCalculator.cs:36-51) per site:
    Z = [E1 E2] [H1 H2]^-1 from the two polarisations' 3-vectors; arrays [..., 3] -> [..., 2, 2]."""
    det = h1[..., 0] * h2[..., 1] - h2[..., 0] * h1[..., 1]
    hi11, hi12, hi21, hi22 = h2[..., 1] / det, -h2[..., 0] / det, -h1[..., 1] / det, h1[..., 0] / det
    z = np.empty(e1.shape[:-1] + (2, 2), dtype=np.complex128)
    z[..., 0, 0] = e1[..., 0] * hi11 + e2[..., 0] * hi21
    z[..., 0, 1] = e1[..., 0] * hi12 + e2[..., 0] * hi22
    z[..., 1, 0] = e1[..., 1] * hi11 + e2[..., 1] * hi21
    z[..., 1, 1] = e1[..., 1] * hi12 + e2[..., 1] * hi22
    return z


def rho_phase(z, omega):
    """PlainTextExporter.CalculateApparentResistivity / ProcessPhase (:360-371, 417-420): rho_a =
    |Z|^2 / (omega mu0), phase = arg Z in radians."""
    return np.abs(z) ** 2 / (omega * xo_mu0()), np.angle(z)


def xo_mu0():
    return lib().xo_mu0()


def contract(nz, green, src):
    src = np.ascontiguousarray(src, dtype=np.complex128).reshape(-1)
    length = src.size // (3 * nz)
    dst = np.zeros_like(src)
    lib().xo_contract(nz, length, *[np.ascontiguousarray(green[c]).reshape(-1) for c in COMPONENTS], src, dst)
    return dst


class Operator:
    """ConvolutionOperator restated (Forward/AtoA/ConvolutionOperator.cs)."""

    def __init__(self, model, green):
        self.model = model
        self._g = {c: np.ascontiguousarray(green[c]).reshape(-1) for c in COMPONENTS}
        self.op = Op()
        rc = lib().xo_op_init(C.byref(self.op), model.nx, model.ny, model.nz, *[self._g[c] for c in COMPONENTS])
        if rc:
            raise MemoryError("xo_op_init")

    def prepare(self, op_type, zeta, zeta0):
        """op_type 'A' or 'Chi0'."""
        self._zeta = np.ascontiguousarray(zeta, dtype=np.complex128).reshape(-1)
        self._zeta0 = np.ascontiguousarray(zeta0, dtype=np.complex128)
        self._dz = np.ascontiguousarray(self.model.thick, dtype=np.float64)
        lib().xo_op_prepare(C.byref(self.op), 0 if op_type == "A" else 1, self._zeta, self._zeta0, self._dz)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
        y = np.zeros_like(x)
        lib().xo_op_apply(C.byref(self.op), x.ctypes.data, y.ctypes.data)
        return y

    def raw_apply(self, xptr, yptr):
        lib().xo_op_apply(C.byref(self.op), xptr, yptr)

    def __del__(self):
        try:
            lib().xo_op_free(C.byref(self.op))
        except Exception:
            pass


def fgmres(matvec, b, x0, m_outer=30, m_inner=10, max_repeat=100, tol=1e-8, backward_check=1):
    """FlexibleGmresWithGmresPreconditioner.Solve(b, x0, x0). matvec(xptr, yptr) works on raw pointers.
    Returns (x, info dict, list of (outer?, iteration, bea))."""
    b = np.ascontiguousarray(b, dtype=np.complex128).reshape(-1)
    x = np.ascontiguousarray(x0, dtype=np.complex128).reshape(-1).copy()
    hist = []

    def _mv(ctx, xp, yp):
        matvec(xp, yp)

    def _it(ctx, outer, it, bea):
        hist.append((outer, it, bea))

    mv = MATVEC(_mv)
    itf = ITER(_it)
    p = GmresParams(b.size, m_outer, m_inner, max_repeat, tol, backward_check)
    r = Result()
    lib().xo_fgmres_solve(C.byref(p), mv, None, itf, None, b, x, C.byref(r))
    info = {"converged": bool(r.converged), "bea": r.bea, "be": r.be, "iterations": r.iterations,
            "inner_iterations": r.inner_iterations, "matvecs": r.matvecs, "dots": r.dots}
    return x, info, hist


def zrotg(a, b):
    c = C.c_double()
    s = CDouble()
    lib().xo_zrotg(CDouble(a.real, a.imag), CDouble(b.real, b.imag), C.byref(c), C.byref(s))
    return c.value, complex(s.re, s.im)


def normal_field(model, omega):
    """E0 per anomaly layer (Mt3DForwardSolver.CalculateNormalFieldFromSource)."""
    return np.array([plane_wave_e(model.section, omega, float(d + t / 2), float(t))
                     for d, t in zip(model.depth, model.thick)], dtype=np.complex128)


def jscattered(model, pol, zeta, zeta0, e0):
    """j_s = (zeta - zeta0) E0 in the [Nx][Ny][3][Nz] layout (ForwardSolver.cs:225-237)."""
    nx, ny, nz = model.nx, model.ny, model.nz
    js = np.zeros((nx, ny, 3, nz), dtype=np.complex128)
    js[:, :, pol, :] = (zeta.reshape(nx, ny, nz) - zeta0[None, None, :]) * e0[None, None, :]
    return js.reshape(-1)


def escattered_jq(model, zeta, zeta0, chi, js):
    nx, ny, nz = model.nx, model.ny, model.nz
    z = zeta.reshape(nx, ny, 1, nz)
    z0 = zeta0.reshape(1, 1, 1, nz)
    chi = chi.reshape(nx, ny, 3, nz)
    js = js.reshape(nx, ny, 3, nz)
    es = (2 * np.sqrt(z0.real + 0j) * chi - js) / (z + np.conj(z0))
    jq = es * (z - z0) + js
    return es.reshape(-1), jq.reshape(-1)
