"""Python host interface of libxtreme_hip.so, mirroring the reference's operator and solver surface.

Names follow the C# classes they stand in for:
  ConvolutionOperator          Extreme.Cartesian/Forward/AtoA/ConvolutionOperator.cs
  AnomalyCurrentFgmresSolver   Extreme.Cartesian/Forward/AnomalyCurrentFgmresSolver.cs
  ForwardSettings              Extreme.Cartesian/Forward/Project/ForwardSettings.cs
  ResultInfo                   Extreme.Fgmres/ResultInfo.cs

Device vectors are torch tensors (complex128, on the GPU) used only as memory; every computation
runs in the HIP library. There is no CPU fallback: without the built library or a GPU, the
constructors raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import cartesian_model as cm

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxtreme_hip.so")
_lib = None

XTR_FP64, XTR_FP32, XTR_MIXED = 0, 1, 2
_PRECISION = {"fp64": XTR_FP64, "fp32": XTR_FP32, "mixed": XTR_MIXED}
XTR_OP_A, XTR_OP_CHI0 = 0, 1


class XtrError(RuntimeError):
    pass


class _Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("depth", C.POINTER(C.c_double)), ("thick", C.POINTER(C.c_double)), ("nlayers", C.c_int),
                ("zero_level", C.c_double), ("layer_thickness", C.POINTER(C.c_double)),
                ("layer_sigma", C.POINTER(C.c_double)), ("omega", C.c_double), ("number_of_hankels", C.c_int),
                ("hank0", C.POINTER(C.c_double)), ("hank1", C.POINTER(C.c_double)), ("ntaps", C.c_int),
                ("precision", C.c_int), ("max_rhs", C.c_int)]


class _Params(C.Structure):
    _fields_ = [("outer_buffer_length", C.c_int), ("inner_buffer_length", C.c_int),
                ("max_repeat_number", C.c_int), ("tolerance", C.c_double), ("backward_error_checking", C.c_int)]


class _Info(C.Structure):
    _fields_ = [("converged", C.c_int), ("arnoldi_backward_error", C.c_double), ("backward_error", C.c_double),
                ("number_of_iterations", C.c_int), ("matvecs", C.c_int64), ("dot_products", C.c_int64),
                ("seconds", C.c_double)]


class _RefineParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("inner_floor", C.c_double), ("max_refinements", C.c_int),
                ("inner", _Params), ("vector_precision", C.c_int)]


class _RefineInfo(C.Structure):
    _fields_ = [("converged", C.c_int), ("residual", C.c_double), ("refinements", C.c_int),
                ("matvecs64", C.c_int64), ("matvecs32", C.c_int64), ("inner_iterations", C.c_int),
                ("seconds", C.c_double)]


ITER_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_double)


def lib():
    """Loads libxtreme_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise XtrError(f"{LIB_PATH} is missing: build it with `python -m extreme_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.xtr_op_create.argtypes = [C.POINTER(_Desc), C.POINTER(C.c_void_p)]
        L.xtr_op_destroy.argtypes = [C.c_void_p]
        L.xtr_op_destroy.restype = None
        L.xtr_op_build_green.argtypes = [C.c_void_p]
        L.xtr_op_set_green.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.xtr_op_get_green.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.xtr_op_prepare.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.xtr_op_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.xtr_op_apply_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.xtr_op_apply_c64_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.xtr_fgmres_solve.argtypes = [C.c_void_p, C.POINTER(_Params), C.c_void_p, C.c_void_p, C.POINTER(_Info),
                                       ITER_FN, C.c_void_p]
        L.xtr_fgmres_solve_device.argtypes = L.xtr_fgmres_solve.argtypes
        L.xtr_fgmres_solve_batch_device.argtypes = [C.c_void_p, C.POINTER(_Params), C.c_int,
                                                    C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                                    C.POINTER(_Info), ITER_FN, C.c_void_p]
        L.xtr_refined_solve_batch_device.argtypes = [C.c_void_p, C.POINTER(_RefineParams), C.c_int,
                                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                                     C.POINTER(_RefineInfo)]
        L.xtr_refined_solve_models_device.argtypes = [C.c_void_p, C.POINTER(_RefineParams), C.c_int,
                                                      C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                                      C.POINTER(C.c_void_p), C.POINTER(_RefineInfo)]
        L.xtr_op_prepare_models.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.xtr_op_apply_models_device.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                                 C.POINTER(C.c_void_p)]
        L.xtr_fgmres_solve_models_device.argtypes = [C.c_void_p, C.POINTER(_Params), C.c_int, C.POINTER(C.c_int),
                                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                                     C.POINTER(_Info), ITER_FN, C.c_void_p]
        L.xtr_op_build_green_atoo.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int]
        L.xtr_op_get_green_atoo.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
        L.xtr_op_jq_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.xtr_op_jq_model_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.xtr_op_level_field_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.xtr_op_level_field_noalloc_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.xtr_op_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.xtr_op_stage_times.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.xtr_op_stage_times_rhs.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.xtr_op_fft_tiles.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.xtr_op_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.xtr_op_green_layout.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.xtr_op_green_mirror_check.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.xtr_last_error.argtypes = [C.c_char_p, C.c_int]
        L.xtr_probe_stream.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p]
        f32p = C.POINTER(C.c_float)
        vpp = C.POINTER(C.c_void_p)
        L.xtr_net_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, vpp]
        L.xtr_net_set_linear.argtypes = [C.c_void_p, C.c_int]
        L.xtr_net_set_features.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.xtr_net_destroy.argtypes = [C.c_void_p]
        L.xtr_net_destroy.restype = None
        L.xtr_net_set_weights.argtypes = [C.c_void_p, f32p, f32p]
        L.xtr_net_get_weights.argtypes = [C.c_void_p, f32p, f32p]
        L.xtr_net_apply_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.xtr_net_gradients_device.argtypes = [C.c_void_p, C.c_int, vpp, vpp, f32p, f32p, C.POINTER(C.c_double)]
        L.xtr_net_train_device.argtypes = [C.c_void_p, C.c_int, vpp, vpp, C.c_float, C.POINTER(C.c_double)]
        L.xtr_net_set_spectral.argtypes = [C.c_void_p, C.c_int]
        L.xtr_net_spectral_size.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.xtr_net_set_spectral_weights.argtypes = [C.c_void_p, f32p]
        L.xtr_net_get_spectral_weights.argtypes = [C.c_void_p, f32p]
        L.xtr_net_spectral_gradient.argtypes = [C.c_void_p, f32p]
        L.xtr_fgmres_solve_learned_device.argtypes = [C.c_void_p, C.POINTER(_Params), C.c_void_p, C.c_void_p,
                                                      C.c_void_p, C.POINTER(_Info)]
        L.xtr_fgmres_solve_harvest_device.argtypes = [C.c_void_p, C.POINTER(_Params), C.c_void_p, C.c_void_p,
                                                      C.POINTER(_Info), C.c_int, vpp, vpp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        buf = C.create_string_buffer(4096)
        lib().xtr_last_error(buf, 4096)
        raise XtrError(buf.value.decode(errors="replace"))


def hankel_table():
    """The 431-tap ndec=40 J0/J1 filter (HankelCoefficients.LoadN40, HankelCoefficients.cs:37,40-65)."""
    t = np.loadtxt(os.path.join(_HERE, "data", "hankel_ndec40.txt"))
    return np.ascontiguousarray(t[:, 0]), np.ascontiguousarray(t[:, 1])


def probe_stream(src, dst, blocks, read_kb, write_kb):
    """Launches the HBM stream probe on torch's current stream: `blocks` blocks of read_kb KB read from
    the device tensor src into LDS as the streaming contraction kernels copy the Green tensor, write_kb
    KB written to dst after each block (xtr_probe_stream). No counterpart in the reference; a
    measurement aid for the read/write mix the contraction streams."""
    torch = _torch()
    for t, kb, what in ((src, read_kb, "src"), (dst, write_kb, "dst")):
        if kb == 0 and what == "dst":
            continue
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
            raise ValueError(f"probe_stream: {what} must be a contiguous GPU tensor")
        if t.numel() * t.element_size() < blocks * kb * 1024:
            raise ValueError(f"probe_stream: {what} holds fewer than blocks * {what}_kb KB")
    stream = torch.cuda.current_stream().cuda_stream
    _check(lib().xtr_probe_stream(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr() if write_kb else 0),
                                  blocks, read_kb, write_kb, C.c_void_p(stream)))


@dataclass
class ForwardSettings:
    """ForwardSettings (Forward/Project/ForwardSettings.cs:6-18); the defaults are the project
    reader's for a project that lists all five values."""
    residual: float = 1e-8
    inner_buffer_length: int = 10
    outer_buffer_length: int = 30
    max_repeats_number: int = 100
    number_of_hankels: int = 10


@dataclass
class RefineInfo:
    """Result of one mixed-precision solve: the FP64 true relative residual it stopped on and the
    applies it spent in each precision."""
    converged: bool
    residual: float
    refinements: int
    matvecs64: int
    matvecs32: int
    inner_iterations: int
    seconds: float

    @staticmethod
    def _from(i) -> "RefineInfo":
        return RefineInfo(bool(i.converged), i.residual, i.refinements, i.matvecs64, i.matvecs32,
                          i.inner_iterations, i.seconds)


@dataclass
class ResultInfo:
    """ResultInfo (Extreme.Fgmres/ResultInfo.cs:4-36) with operation counts and wall time."""
    convergence_achieved: bool
    arnoldi_backward_error: float
    backward_error: float
    number_of_iterations: int
    matvecs: int
    dot_products: int
    seconds: float

    @staticmethod
    def _from(i: _Info) -> "ResultInfo":
        return ResultInfo(bool(i.converged), i.arnoldi_backward_error, i.backward_error, i.number_of_iterations,
                          i.matvecs, i.dot_products, i.seconds)


def _dptr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _device_vectors(ts, n, what="vector"):
    """Checks tensors handed to a device entry point: the kernels read and write n complex128 values
    behind each bare pointer and cannot check them, so a CPU tensor, another dtype, a strided view or a
    short vector would make them access memory out of bounds. Raises ValueError instead."""
    torch = _torch()
    for t in ts:
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{what} must be a torch tensor on the GPU, got {type(t).__name__}")
        if not t.is_cuda:
            raise ValueError(f"{what} must be on the GPU, got a tensor on {t.device}")
        if t.dtype != torch.complex128:
            raise ValueError(f"{what} must be complex128, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
        if t.numel() != n:
            raise ValueError(f"{what} must hold {n} values, got {t.numel()}")


def _host_vector(a, n, what="vector"):
    if not isinstance(a, np.ndarray) or a.dtype != np.complex128 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{what} must be a C-contiguous complex128 numpy array")
    if a.size != n:
        raise ValueError(f"{what} must hold {n} values, got {a.size}")


def _ptr(t):
    """Device pointer of a torch tensor or host pointer of a numpy array."""
    if hasattr(t, "data_ptr"):
        return C.c_void_p(t.data_ptr())
    return C.c_void_p(t.ctypes.data)


class ConvolutionOperator:
    """The CIE operator A x = x - R x - b IFFT(G FFT(f R x)) and the Chi0 operator on the GPU
    (ConvolutionOperator.cs:45-327), for one model at one frequency."""

    COMPONENTS = ("xx", "xy", "xz", "yy", "yz", "zz")

    def __init__(self, model: cm.CartesianModel, omega: float, settings: ForwardSettings | None = None,
                 precision: str = "fp64", max_rhs: int = 2):
        self.model = model
        self.omega = float(omega)
        self.settings = settings or ForwardSettings()
        self.precision = precision
        self.max_rhs = max_rhs
        h0, h1 = hankel_table()
        keep = []
        arrs = {}
        for name, a in (("depth", model.depth), ("thick", model.thick), ("lt", model.section.thickness),
                        ("ls", model.section.sigma), ("h0", h0), ("h1", h1)):
            arr, p = _dptr(a)
            keep.append(arr)
            arrs[name] = p
        d = _Desc(model.nx, model.ny, model.nz, float(model.dx), float(model.dy), arrs["depth"], arrs["thick"],
                  model.section.nlayers, float(model.section.zero_level), arrs["lt"], arrs["ls"], self.omega,
                  int(self.settings.number_of_hankels), arrs["h0"], arrs["h1"], len(h0),
                  _PRECISION[precision], int(max_rhs))
        h = C.c_void_p()
        _check(lib().xtr_op_create(C.byref(d), C.byref(h)))
        self._h = h
        self.n = model.n
        self.operator_type = None

    def close(self):
        if getattr(self, "_h", None):
            lib().xtr_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- Green tensor -------------------------------------------------------------------------
    def calculate_green_tensor(self):
        """AtoAGreenTensorCalculatorComponent.CalculateGreenTensor on the device (FP64)."""
        _check(lib().xtr_op_build_green(self._h))

    def set_green_tensor(self, green: dict):
        """Uploads a Green tensor in the reference layout (component -> [2Nx*2Ny, P] complex128)."""
        arrs = [np.ascontiguousarray(green[c], dtype=np.complex128) for c in self.COMPONENTS]
        _check(lib().xtr_op_set_green(self._h, *[C.c_void_p(a.ctypes.data) for a in arrs]))

    def green_tensor(self) -> dict:
        m = self.model
        L = 4 * m.nx * m.ny
        S = m.nz * (m.nz + 1) // 2
        out = {c: np.zeros((L, m.nz * m.nz if c in ("xz", "yz") else S), dtype=np.complex128)
               for c in self.COMPONENTS}
        _check(lib().xtr_op_get_green(self._h, *[C.c_void_p(out[c].ctypes.data) for c in self.COMPONENTS]))
        return out

    # --- operator -----------------------------------------------------------------------------
    def prepare_operator(self, operator_type: str, zeta=None):
        """PrepareOperator(greenTensor, OperatorType.A | OperatorType.Chi0) (ConvolutionOperator.cs:64-78)."""
        if zeta is None:
            zeta = self.model.zeta(self.omega)
        z = np.ascontiguousarray(np.asarray(zeta, dtype=np.complex128).reshape(-1))
        if z.size != self.model.nx * self.model.ny * self.model.nz:
            raise ValueError("zeta must hold Nx*Ny*Nz values")
        t = {"A": XTR_OP_A, "Chi0": XTR_OP_CHI0}[operator_type]
        _check(lib().xtr_op_prepare(self._h, t, C.c_void_p(z.ctypes.data)))
        self.operator_type = operator_type

    def prepare_models(self, operator_type: str, zetas):
        """PrepareOperator for several conductivity models that share the grid, background and
        frequency (one Green tensor): zetas[m] is model m's [Nx][Ny][Nz] complex conductivity."""
        arrs = [np.ascontiguousarray(np.asarray(z, dtype=np.complex128).reshape(-1)) for z in zetas]
        n = self.model.nx * self.model.ny * self.model.nz
        if any(a.size != n for a in arrs):
            raise ValueError("each zeta must hold Nx*Ny*Nz values")
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        t = {"A": XTR_OP_A, "Chi0": XTR_OP_CHI0}[operator_type]
        _check(lib().xtr_op_prepare_models(self._h, t, len(arrs), ptrs))
        self.operator_type = operator_type

    def apply_models(self, xs, models, ys=None):
        """One apply of device vectors xs[r] on models[r]: the Green tensor is streamed once."""
        n = len(xs)
        if ys is None:
            _device_vectors(xs, self.n, "x")
            ys = [x.new_empty(x.shape) for x in xs]
        if len(models) != n or len(ys) != n:
            raise ValueError("xs, ys and models must have the same length")
        _device_vectors(list(xs) + list(ys), self.n, "vector")
        mi = (C.c_int * n)(*[int(m) for m in models])
        xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        yp = (C.c_void_p * n)(*[y.data_ptr() for y in ys])
        _check(lib().xtr_op_apply_models_device(self._h, n, mi, xp, yp))
        return ys

    def apply(self, x, y=None):
        """Apply(input, output): numpy arrays (host) or torch GPU tensors (device), n or [nrhs, n]."""
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        if hasattr(x, "data_ptr"):
            _device_vectors([x], nrhs * self.n, "x")
            if y is None:
                y = x.new_empty(x.shape)
            _device_vectors([y], nrhs * self.n, "y")
            _check(lib().xtr_op_apply_device(self._h, _ptr(x), _ptr(y), nrhs))
            return y
        x = np.ascontiguousarray(x, dtype=np.complex128)
        _host_vector(x, nrhs * self.n, "x")
        if y is None:
            y = np.empty_like(x)
        _host_vector(y, nrhs * self.n, "y")
        _check(lib().xtr_op_apply(self._h, _ptr(x), _ptr(y), nrhs))
        return y

    def apply_c64(self, x, y=None):
        """The apply in complex64 on complex64 device tensors (n or [nrhs, n]); the operator must be
        created with precision "fp32" or "mixed"."""
        torch = _torch()
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        for t, what in ((x, "x"), (y, "y")):
            if t is None:
                continue
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex64 and t.is_contiguous()
                    and t.numel() == nrhs * self.n):
                raise ValueError(f"{what} must be a contiguous complex64 GPU tensor of {nrhs * self.n} values")
        if y is None:
            y = x.new_empty(x.shape)
        _check(lib().xtr_op_apply_c64_device(self._h, _ptr(x), _ptr(y), nrhs))
        return y

    # --- observation level (AtoO) ----------------------------------------------------------
    AO_COMPONENTS = ("xx", "yy", "zz", "xy", "yx", "xz", "yz", "zx", "zy")
    AO_ELECTRIC = ("xx", "yy", "zz", "xy", "xz", "yz", "zx", "zy")
    AO_MAGNETIC = ("xx", "xy", "yx", "xz", "yz", "zx", "zy")

    def calculate_green_atoo(self, level_z: float = 0.0, electric: bool = True, x_shift: float = 0.0,
                             y_shift: float = 0.0):
        """AtoOCalculator.CalculateGreenTensorAtoOLevelElectric / ...Magnetic (AtoOCalculator.cs:124-160):
        the tensor of the level at depth level_z whose sites are the cell centres shifted by
        (x_shift, y_shift) metres, assembled and FFT'd on the GPU and kept in the handle."""
        _check(lib().xtr_op_build_green_atoo(self._h, float(level_z), float(x_shift), float(y_shift),
                                             1 if electric else 0))

    def green_atoo(self, electric: bool = True) -> dict:
        """The level tensor copied back, component -> [2Nx*2Ny, Nz]."""
        m = self.model
        L = 4 * m.nx * m.ny
        comps = self.AO_ELECTRIC if electric else self.AO_MAGNETIC
        out = {c: np.zeros((L, m.nz), dtype=np.complex128) for c in comps}
        ptrs = (C.c_void_p * 9)(*[out[c].ctypes.data if c in out else None for c in self.AO_COMPONENTS])
        _check(lib().xtr_op_get_green_atoo(self._h, 1 if electric else 0, ptrs))
        return out

    def jq(self, chi, js, out=None, model=None):
        """ForwardSolver.CalculateJqFrom (ForwardSolver.cs:364-389, 424-436) on device tensors; the
        operator must be prepared as "A". model names the model of an operator prepared by
        prepare_models (required when it holds more than one)."""
        _device_vectors([chi, js], self.n, "chi and js")
        if out is None:
            out = chi.new_empty(chi.shape)
        _device_vectors([out], self.n, "out")
        if model is None:
            _check(lib().xtr_op_jq_device(self._h, _ptr(chi), _ptr(js), _ptr(out)))
        else:
            _check(lib().xtr_op_jq_model_device(self._h, int(model), _ptr(chi), _ptr(js), _ptr(out)))
        return out

    def level_field(self, jq, electric: bool = True, out=None):
        """AtoOCalculator.CalculateAnomalyFieldE/H at the level (AtoOCalculator.cs:63-99): device
        tensor [Nx, Ny, 3]."""
        m = self.model
        _device_vectors([jq], self.n, "jq")
        if out is None:
            out = jq.new_empty((m.nx, m.ny, 3))
        _device_vectors([out], 3 * m.nx * m.ny, "out")
        _check(lib().xtr_op_level_field_device(self._h, 1 if electric else 0, _ptr(jq), _ptr(out)))
        return out

    def set_profiling(self, on: bool):
        """Starts (and resets) or stops per-stage HIP-event timing of every apply."""
        _check(lib().xtr_op_set_profiling(self._h, 1 if on else 0))

    STAGES = ("fwd_y", "fwd_x", "contract", "inv_x", "inv_y")

    def stage_times(self, nrhs: int = 0):
        """(summed milliseconds per stage, number of applies) since profiling was switched on, over
        the applies of nrhs right-hand sides (0: all)."""
        out = (C.c_double * 5)()
        n = C.c_int64()
        _check(lib().xtr_op_stage_times_rhs(self._h, int(nrhs), out, C.byref(n)))
        return dict(zip(self.STAGES, list(out))), n.value

    FFT_TILES = ("default", "narrow", "wide", "narrow512", "reg", "reg_alt")

    def fft_tiles(self):
        """Column tile of each FFT pass (fwd_y, fwd_x, inv_x, inv_y) in the latest apply: chosen
        by measurement once per process and shape (or fixed by XTR_FFT_TILES); None before the
        first apply."""
        out = (C.c_int * 4)()
        _check(lib().xtr_op_fft_tiles(self._h, out))
        return None if out[0] < 0 else dict(zip(("fwd_y", "fwd_x", "inv_x", "inv_y"),
                                                (self.FFT_TILES[t] for t in out)))

    def sizes(self):
        n, gb, bb = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().xtr_op_sizes(self._h, C.byref(n), C.byref(gb), C.byref(bb)))
        return n.value, gb.value, bb.value

    def green_mirror_check(self):
        """{component: (max mirror deviation, max |G|)} of a handle holding every block (layout "full")."""
        out = (C.c_double * 12)()
        _check(lib().xtr_op_green_mirror_check(self._h, out))
        return {c: (out[2 * k], out[2 * k + 1]) for k, c in enumerate(("xx", "xy", "yy", "zz", "xz", "yz"))}

    def green_layout(self):
        """("folded", (Nx+1)(Ny+1)) when the device keeps only the blocks kx <= Nx, ky <= Ny (the
        mirror images follow by sign), else ("full", 4 Nx Ny)."""
        f, nb = C.c_int(), C.c_int64()
        _check(lib().xtr_op_green_layout(self._h, C.byref(f), C.byref(nb)))
        return ("folded" if f.value else "full"), nb.value


class AnomalyCurrentFgmresSolver:
    """FGMRES(outer) with an inner GMRES right preconditioner on the GPU
    (AnomalyCurrentFgmresSolver.cs:31-85 + FlexibleGmresWithGmresPreconditioner.cs)."""

    def __init__(self, operator: ConvolutionOperator, settings: ForwardSettings | None = None,
                 backward_error_checking: int = 1):
        self.op = operator
        s = settings or operator.settings
        self.params = _Params(int(s.outer_buffer_length), int(s.inner_buffer_length), int(s.max_repeats_number),
                              float(s.residual), int(backward_error_checking))
        self.history = []

    def _cb(self):
        hist = self.history

        def f(user, rhs, it, bea):
            hist.append((rhs, it, bea))

        return ITER_FN(f)

    def solve(self, b, x) -> ResultInfo:
        """Solves A x = b; x holds the initial guess (the reference passes x0 = chi0)."""
        info = _Info()
        cb = self._cb()
        if hasattr(b, "data_ptr"):
            _device_vectors([b, x], self.op.n, "b and x")
            _check(lib().xtr_fgmres_solve_device(self.op._h, C.byref(self.params), _ptr(b), _ptr(x),
                                                 C.byref(info), cb, None))
        else:
            _host_vector(b, self.op.n, "b")
            _host_vector(x, self.op.n, "x")
            _check(lib().xtr_fgmres_solve(self.op._h, C.byref(self.params), _ptr(b), _ptr(x), C.byref(info),
                                          cb, None))
        return ResultInfo._from(info)

    def solve_learned(self, b, x, net: "LearnedPreconditioner") -> ResultInfo:
        """FGMRES with the learned network as the right preconditioner in place of the inner GMRES
        (device tensors; x holds the initial guess)."""
        _device_vectors([b, x], self.op.n, "b and x")
        if net.shape != (self.op.model.nx, self.op.model.ny, self.op.model.nz):
            raise ValueError("the network's grid differs from the operator's")
        info = _Info()
        _check(lib().xtr_fgmres_solve_learned_device(self.op._h, C.byref(self.params), net._h, _ptr(b), _ptr(x),
                                                     C.byref(info)))
        return ResultInfo._from(info)

    def solve_harvest(self, b, x, capacity: int):
        """solve() on device tensors that also returns the first `capacity` pairs (v_j, z_j) of the
        flexible preconditioner: the Krylov vector and the inner GMRES's result for it."""
        torch = _torch()
        _device_vectors([b, x], self.op.n, "b and x")
        vs = [torch.empty_like(b) for _ in range(capacity)]
        zs = [torch.empty_like(b) for _ in range(capacity)]
        vp = (C.c_void_p * max(capacity, 1))(*[v.data_ptr() for v in vs])
        zp = (C.c_void_p * max(capacity, 1))(*[z.data_ptr() for z in zs])
        info, cnt = _Info(), C.c_int()
        _check(lib().xtr_fgmres_solve_harvest_device(self.op._h, C.byref(self.params), _ptr(b), _ptr(x), C.byref(info),
                                                     capacity, vp, zp, C.byref(cnt)))
        return ResultInfo._from(info), vs[:cnt.value], zs[:cnt.value]

    def solve_batch(self, bs, xs) -> list:
        """Side-by-side solves sharing every operator apply (device tensors)."""
        n = len(bs)
        if len(xs) != n:
            raise ValueError("bs and xs must have the same length")
        _device_vectors(list(bs) + list(xs), self.op.n, "b and x")
        infos = (_Info * n)()
        bp = (C.c_void_p * n)(*[b.data_ptr() for b in bs])
        xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        cb = self._cb()
        _check(lib().xtr_fgmres_solve_batch_device(self.op._h, C.byref(self.params), n, bp, xp, infos, cb, None))
        return [ResultInfo._from(infos[i]) for i in range(n)]

    def solve_models_batch(self, bs, xs, models) -> list:
        """Side-by-side solves of right-hand sides on different models that share the Green tensor
        (operator prepared with prepare_models("A", ...)); device tensors."""
        n = len(bs)
        if len(xs) != n or len(models) != n:
            raise ValueError("bs, xs and models must have the same length")
        _device_vectors(list(bs) + list(xs), self.op.n, "b and x")
        infos = (_Info * n)()
        mi = (C.c_int * n)(*[int(m) for m in models])
        bp = (C.c_void_p * n)(*[b.data_ptr() for b in bs])
        xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        cb = self._cb()
        _check(lib().xtr_fgmres_solve_models_device(self.op._h, C.byref(self.params), n, mi, bp, xp, infos, cb,
                                                    None))
        return [ResultInfo._from(infos[i]) for i in range(n)]

    def solve_refined_batch(self, bs, xs, tolerance=None, inner_floor=1e-6, max_refinements=10,
                            vectors: str = "fp32", models=None) -> list:
        """Mixed precision (operator created with precision="mixed"): FP64 iterative refinement
        around complex64 FGMRES corrections, stopping on the FP64 true residual. The corrections'
        Krylov vectors are complex64 (vectors="fp32") or complex128 ("fp64"); their dots, norms and
        Hessenberg matrix are FP64 either way. models[r] (operator prepared with
        prepare_models("A", ...)) names right-hand side r's model. Device tensors."""
        n = len(bs)
        if len(xs) != n or (models is not None and len(models) != n):
            raise ValueError("bs, xs and models must have the same length")
        _device_vectors(list(bs) + list(xs), self.op.n, "b and x")
        infos = (_RefineInfo * n)()
        bp = (C.c_void_p * n)(*[b.data_ptr() for b in bs])
        xp = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        tol = float(self.params.tolerance if tolerance is None else tolerance)
        rp = _RefineParams(tol, float(inner_floor), int(max_refinements), self.params, _PRECISION[vectors])
        if models is None:
            _check(lib().xtr_refined_solve_batch_device(self.op._h, C.byref(rp), n, bp, xp, infos))
        else:
            mi = (C.c_int * n)(*[int(m) for m in models])
            _check(lib().xtr_refined_solve_models_device(self.op._h, C.byref(rp), n, mi, bp, xp, infos))
        return [RefineInfo._from(infos[i]) for i in range(n)]


def _torch():
    import torch
    return torch


class LearnedPreconditioner:
    """A learned right preconditioner for the flexible FGMRES (no counterpart in the reference):
    z = s (v/s + N(v/s)), s = rms(v), N a two-layer 3 x 3 x 3 convolution network over the grid's
    cells with channels (Re, Im) x (x, y, z), BF16 MFMA convolutions, FP32 Adam. Weights:
    w1[t, ci, co] (27, 8, 16), w2[t, ci, co] (27, 16, 8); tap t = offset (t//9 - 1, t//3 % 3 - 1,
    t % 3 - 1) in (i, j, k); channel 2c + p = part p (Re, Im) of component c; channels 6, 7 the cell's
    features (log10 sigma, log10 f) once set_features is called, else zero.
    linear=True drops the hidden layer's relu: the preconditioner is then a fixed linear map."""

    def __init__(self, nx: int, ny: int, nz: int, seed: int = 1, linear: bool = False, modes: int = 0):
        h = C.c_void_p()
        _check(lib().xtr_net_create(nx, ny, nz, C.c_uint64(seed), C.byref(h)))
        self._h = h
        self.shape = (nx, ny, nz)
        self.linear = bool(linear)
        self.modes = 0
        if linear:
            _check(lib().xtr_net_set_linear(self._h, 1))
        if modes:
            self.set_spectral(modes)

    # --- spectral branch: per-lateral-mode mixing of the 3 nz complex channels, modes |kx|, |ky| <= K
    def set_spectral(self, modes: int):
        """Adds (modes > 0, zero weights) or removes (0) the spectral branch."""
        _check(lib().xtr_net_set_spectral(self._h, int(modes)))
        self.modes = int(modes)

    def _spec_shape(self):
        nm = (2 * self.modes + 1) ** 2
        q = 3 * self.shape[2]
        return (2, nm, q, q)

    def parameter_count(self) -> int:
        """Trainable FP32 parameters: the two convolutions and the spectral branch."""
        k, f = C.c_int(), C.c_int64()
        _check(lib().xtr_net_spectral_size(self._h, C.byref(k), C.byref(f)))
        return 27 * 8 * 16 + 27 * 16 * 8 + f.value

    def spectral_weights(self):
        """(W0, W1), complex64 [NM, 3nz, 3nz] each."""
        w = np.zeros(self._spec_shape() + (2,), dtype=np.float32)
        _check(lib().xtr_net_get_spectral_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float))))
        c = w[..., 0] + 1j * w[..., 1]
        return c[0].astype(np.complex64), c[1].astype(np.complex64)

    def set_spectral_weights(self, w0, w1):
        shp = self._spec_shape()[1:]
        w = np.zeros((2,) + shp + (2,), dtype=np.float32)
        for k, a in enumerate((w0, w1)):
            a = np.asarray(a).reshape(shp)
            w[k, ..., 0] = a.real
            w[k, ..., 1] = a.imag
        _check(lib().xtr_net_set_spectral_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float))))

    def spectral_gradient(self):
        """Gradient of the last gradients()/train_step() call: (dL/dW0, dL/dW1) as complex [NM, 3nz, 3nz]
        (real part: derivative by the real part of the weight, imaginary part: by its imaginary part)."""
        g = np.zeros(self._spec_shape() + (2,), dtype=np.float32)
        _check(lib().xtr_net_spectral_gradient(self._h, g.ctypes.data_as(C.POINTER(C.c_float))))
        c = g[..., 0] + 1j * g[..., 1]
        return c[0], c[1]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().xtr_net_destroy(self._h)
            self._h = None

    def set_features(self, sigma=None, frequency=None):
        """Input channels 6 and 7 of every cell: log10 of the cell's conductivity sigma (S/m, shape
        (nx, ny, nz)) and log10 of the frequency (Hz); None clears them to zeros."""
        if sigma is None:
            _check(lib().xtr_net_set_features(self._h, None))
            return
        sig = np.asarray(sigma, dtype=np.float64).reshape(self.shape)
        feat = np.empty(sig.shape + (2,), dtype=np.float32)
        feat[..., 0] = np.log10(sig)
        feat[..., 1] = np.log10(frequency)
        feat = np.ascontiguousarray(feat)
        _check(lib().xtr_net_set_features(self._h, feat.ctypes.data_as(C.POINTER(C.c_float))))

    def weights(self):
        w1 = np.zeros((27, 8, 16), dtype=np.float32)
        w2 = np.zeros((27, 16, 8), dtype=np.float32)
        f = C.POINTER(C.c_float)
        _check(lib().xtr_net_get_weights(self._h, w1.ctypes.data_as(f), w2.ctypes.data_as(f)))
        return w1, w2

    def set_weights(self, w1, w2):
        w1 = np.ascontiguousarray(w1, dtype=np.float32).reshape(27, 8, 16)
        w2 = np.ascontiguousarray(w2, dtype=np.float32).reshape(27, 16, 8)
        f = C.POINTER(C.c_float)
        _check(lib().xtr_net_set_weights(self._h, w1.ctypes.data_as(f), w2.ctypes.data_as(f)))

    def apply(self, v, out=None):
        """z = preconditioner(v) for a complex128 device vector of 3 nx ny nz values."""
        n = 3 * self.shape[0] * self.shape[1] * self.shape[2]
        _device_vectors([v], n, "v")
        if out is None:
            out = v.new_empty(v.shape)
        _device_vectors([out], n, "out")
        _check(lib().xtr_net_apply_device(self._h, _ptr(v), _ptr(out)))
        return out

    def _check_pairs(self, vs, zs):
        if len(vs) != len(zs) or not vs:
            raise ValueError("vs and zs must be non-empty and of the same length")
        _device_vectors(list(vs) + list(zs), 3 * self.shape[0] * self.shape[1] * self.shape[2], "pair vector")

    @staticmethod
    def _ptrs(ts):
        return (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def gradients(self, vs, zs):
        """(dL/dw1, dL/dw2, L) for the mean relative error of the correction over the pairs."""
        self._check_pairs(vs, zs)
        g1 = np.zeros((27, 8, 16), dtype=np.float32)
        g2 = np.zeros((27, 16, 8), dtype=np.float32)
        loss = C.c_double()
        f = C.POINTER(C.c_float)
        _check(lib().xtr_net_gradients_device(self._h, len(vs), self._ptrs(vs), self._ptrs(zs), g1.ctypes.data_as(f),
                                              g2.ctypes.data_as(f), C.byref(loss)))
        return g1, g2, loss.value

    def train_step(self, vs, zs, lr: float = 1e-3) -> float:
        """One Adam step on the pairs; returns the loss before the step."""
        self._check_pairs(vs, zs)
        loss = C.c_double()
        _check(lib().xtr_net_train_device(self._h, len(vs), self._ptrs(vs), self._ptrs(zs), C.c_float(lr),
                                          C.byref(loss)))
        return loss.value
