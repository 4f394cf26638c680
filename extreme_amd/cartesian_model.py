"""Cartesian model description and the seeded synthetic models of the benchmark configurations.

A model is what extrEMe's CartesianModel holds (Extreme.Cartesian/Model/CartesianModel.cs,
Extreme.Core/Model/Section1D.cs, Extreme.Core/Model/AnomalyLayer.cs): a 1-D background
section of layers (layer 0 the air half-space, the last layer the lower half-space, both
with thickness 0, as in EndUser/ExtremeMt/Examples/COMMEMI3D2/commemi.xmodel:4-10) and an
anomalous domain of Nx x Ny x Nz cells whose layers must each lie inside one finite
background layer (Extreme.Cartesian/Model/ModelUtils.cs:35-52).

The synthetic inputs are generated with SplitMix64, seeded 0x5EED0000 + 100*config + model,
so that any other implementation can rebuild the same models.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MU0 = 4.0 * math.pi * 1.0e-07                    # OmegaModelUtils.cs:12
EPS0 = 1.0 / (MU0 * 299792458.0 * 299792458.0)   # OmegaModelUtils.cs:13


def depth_um(v: float) -> int:
    """A depth or thickness in whole micrometres. The reference holds them as System.Decimal
    (Extreme.Core/Model/Layer1D.cs), so its layer lookups add and compare exactly; on doubles
    2.2 + 99.9 lands above 102.1. Integers of micrometres are exact for any depth written with at
    most six decimals."""
    return int(round(float(v) * 1e6))


def omega_of(frequency: float) -> float:
    """OmegaModelUtils.FrequencyToOmega (OmegaModelUtils.cs:55-58)."""
    return 2 * math.pi * frequency


def zeta_of(sigma, omega: float):
    """zeta = sigma - i*omega*eps0 (OmegaModelUtils.cs:41-42)."""
    return np.asarray(sigma, dtype=np.float64) - 1j * (omega * EPS0)


@dataclass
class Section1D:
    zero_level: float
    thickness: list
    sigma: list

    @property
    def nlayers(self) -> int:
        return len(self.thickness)

    def layer_depth(self, index: int) -> float:
        """OmegaModel.GetLayerDepth (Model/Omega/OmegaModel.cs:27-38)."""
        return self.zero_level + float(sum(self.thickness[:index]))

    def background_layer(self, start: float, end: float) -> int:
        """ModelUtils.FindCorrespondingBackgroundLayerIndex (Model/ModelUtils.cs:35-52): the first layer
        whose closed interval holds [start, end], compared in exact micrometres (depth_um)."""
        ss = depth_um(self.zero_level)
        a, b = depth_um(start), depth_um(end)
        for i, t in enumerate(self.thickness):
            es = ss + depth_um(t)
            if a >= ss and b <= es:
                return i
            ss = es
        raise ValueError(f"anomaly layer [{start}, {end}] is not inside one background layer")

    def layer_at_depth(self, depth: float) -> int:
        """ScalarMathUtils.DetermineCorrespondingBackgroundLayer, half-open [d1, d2)
        (ScalarMathUtils.cs:15-31), in exact micrometres (depth_um)."""
        prev = depth_um(self.zero_level)
        z = depth_um(depth)
        for k, t in enumerate(self.thickness):
            d2 = prev + depth_um(t)
            if prev <= z < d2:
                return k
            prev = d2
        raise ValueError("depth is not within the 1-D section")

    def zeta(self, omega: float) -> np.ndarray:
        return zeta_of(self.sigma, omega)


@dataclass
class CartesianModel:
    nx: int
    ny: int
    dx: float
    dy: float
    depth: np.ndarray        # [nz] top of each anomaly layer
    thick: np.ndarray        # [nz]
    sigma: np.ndarray        # [nx, ny, nz] anomaly conductivity, S/m
    section: Section1D
    name: str = ""
    frequencies: list = field(default_factory=list)

    @property
    def nz(self) -> int:
        return len(self.depth)

    @property
    def n(self) -> int:
        return 3 * self.nx * self.ny * self.nz

    def background_indices(self) -> list:
        return [self.section.background_layer(float(d), float(d + t)) for d, t in zip(self.depth, self.thick)]

    def zeta(self, omega: float) -> np.ndarray:
        return zeta_of(self.sigma, omega)

    def zeta0(self, omega: float) -> np.ndarray:
        """Background zeta of the layer holding each anomaly layer (ConvolutionOperator.GetBackgroundZeta :80-88)."""
        z = self.section.zeta(omega)
        return np.array([z[i] for i in self.background_indices()], dtype=np.complex128)


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the documented generator of the synthetic models."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] inclusive."""
        return lo + min(hi - lo, int(self.uniform() * (hi - lo + 1)))


def seed_for(config: int, model: int = 0) -> int:
    return 0x5EED0000 + 100 * config + model


def _faces_geometric(nz: int, dz0: float, ratio: float, snaps=()):
    dz = [dz0 * ratio ** (k / (nz - 1)) for k in range(nz)]
    faces = [0.0]
    for d in dz:
        faces.append(faces[-1] + d)
    for s in snaps:
        k = min(range(1, nz), key=lambda i: abs(faces[i] - s))
        faces[k] = float(s)
    return faces


def _paint_boxes(rng: SplitMix64, sigma: np.ndarray, nboxes: int, side_lo: int, side_hi: int,
                 log10_rho_lo: float, log10_rho_hi: float):
    nx, ny, nz = sigma.shape
    for _ in range(nboxes):
        sx = min(rng.randint(side_lo, side_hi), nx)
        sy = min(rng.randint(side_lo, side_hi), ny)
        sz = min(rng.randint(side_lo, side_hi), nz)
        x0 = rng.randint(0, nx - sx)
        y0 = rng.randint(0, ny - sy)
        z0 = rng.randint(0, nz - sz)
        lr = log10_rho_lo + (log10_rho_hi - log10_rho_lo) * rng.uniform()
        sigma[x0:x0 + sx, y0:y0 + sy, z0:z0 + sz] = 10.0 ** (-lr)


def _three_layer_section(bottom: float):
    """Air | 100 ohm-m to 2 km | 10 ohm-m to 10 km | 1000 ohm-m, extended by a same-sigma
    slab down to `bottom` when the domain reaches the lower half-space
    (the split of OneBlockModelCreater.cs:77-85)."""
    thick = [0.0, 2000.0, 8000.0]
    sig = [0.0, 1e-2, 1e-1]
    if bottom > 10000.0:
        thick.append(bottom - 10000.0)
        sig.append(1e-3)
    thick.append(0.0)
    sig.append(1e-3)
    return Section1D(0.0, thick, sig)


def config_c1() -> CartesianModel:
    """Plain method, small: 32x32x8 cells, dx=dy=250 m, dz=125 m over 0-1 km, 100 ohm-m half-space
    below air, one 1 ohm-m 8x8x4 block at the centre from 250 to 750 m, f = 0.1 Hz."""
    nx = ny = 32
    nz = 8
    depth = np.arange(nz) * 125.0
    thick = np.full(nz, 125.0)
    sigma = np.full((nx, ny, nz), 1e-2)
    sigma[12:20, 12:20, 2:6] = 1.0
    sec = Section1D(0.0, [0.0, 1000.0, 0.0], [0.0, 1e-2, 1e-2])
    return CartesianModel(nx, ny, 250.0, 250.0, depth, thick, sigma, sec, "C1", [0.1])


def config_c2(model_index: int = 0) -> CartesianModel:
    """Mixed precision vs FP64: 128x128x32, dx=dy=200 m, dz growing 50 -> 400 m geometric with the
    face nearest 2 km snapped to it, the 3-layer background and 20 seeded random boxes
    (sides 4-32 cells, log-uniform 1-1000 ohm-m); 8 frequencies 1e-3 .. 10 Hz."""
    nx = ny = 128
    nz = 32
    faces = _faces_geometric(nz, 50.0, 8.0, snaps=(2000.0,))
    depth = np.array(faces[:-1])
    thick = np.diff(np.array(faces))
    sec = _three_layer_section(faces[-1])
    sigma = np.empty((nx, ny, nz))
    for k in range(nz):
        sigma[:, :, k] = sec.sigma[sec.background_layer(depth[k], depth[k] + thick[k])]
    _paint_boxes(SplitMix64(seed_for(2, model_index)), sigma, 20, 4, 32, 0.0, 3.0)
    freqs = [10.0 ** (-3 + 4 * k / 7) for k in range(8)]
    return CartesianModel(nx, ny, 200.0, 200.0, depth, thick, sigma, sec, "C2", freqs)


def config_c3(model_index: int = 0) -> CartesianModel:
    """Batched, uneven sizes: model k of 256, Nx and Ny drawn from {32, 48, 64, 96} and Nz from
    {8, 12, 16, 24}, dx = dy = 250 m, dz = 100 m (2 km is a cell face), the 3-layer background and
    20 seeded boxes as in the 128x128x32 model (sides 4-32 cells, log-uniform 1-1000 ohm-m); two
    distinct frequencies of {0.01, 0.1, 1} Hz per model."""
    rng = SplitMix64(seed_for(3, model_index))
    sizes, depths = (32, 48, 64, 96), (8, 12, 16, 24)
    nx = sizes[rng.randint(0, 3)]
    ny = sizes[rng.randint(0, 3)]
    nz = depths[rng.randint(0, 3)]
    skip = rng.randint(0, 2)                        # the frequency of the three not used
    freqs = [f for k, f in enumerate((0.01, 0.1, 1.0)) if k != skip]
    depth = np.arange(nz) * 100.0
    thick = np.full(nz, 100.0)
    sec = _three_layer_section(nz * 100.0)
    sigma = np.empty((nx, ny, nz))
    for k in range(nz):
        sigma[:, :, k] = sec.sigma[sec.background_layer(depth[k], depth[k] + thick[k])]
    _paint_boxes(rng, sigma, 20, 4, 32, 0.0, 3.0)
    return CartesianModel(nx, ny, 250.0, 250.0, depth, thick, sigma, sec, "C3", freqs)


def config_c4(model_index: int = 0) -> CartesianModel:
    """Learned-preconditioner training: model k of 4096, 64x64x16 cells, dx = dy = 250 m, dz = 100 m,
    5-30 seeded boxes (sides 4-32 cells, log-uniform 1-1000 ohm-m) over the 3-layer background;
    frequencies 0.01, 0.1, 1 and 10 Hz."""
    nx = ny = 64
    nz = 16
    rng = SplitMix64(seed_for(4, model_index))
    depth = np.arange(nz) * 100.0
    thick = np.full(nz, 100.0)
    sec = _three_layer_section(nz * 100.0)
    sigma = np.empty((nx, ny, nz))
    for k in range(nz):
        sigma[:, :, k] = sec.sigma[sec.background_layer(depth[k], depth[k] + thick[k])]
    _paint_boxes(rng, sigma, rng.randint(5, 30), 4, 32, 0.0, 3.0)
    return CartesianModel(nx, ny, 250.0, 250.0, depth, thick, sigma, sec, "C4", [0.01, 0.1, 1.0, 10.0])


def config_c5(model_index: int = 0) -> CartesianModel:
    """Large: 256x256x64, dx=dy=100 m, dz 25 -> 800 m geometric with faces snapped to 2 and 10 km,
    the 3-layer background plus 60 seeded boxes (log-uniform 0.3-3000 ohm-m); 12 frequencies."""
    nx = ny = 256
    nz = 64
    faces = _faces_geometric(nz, 25.0, 32.0, snaps=(2000.0, 10000.0))
    depth = np.array(faces[:-1])
    thick = np.diff(np.array(faces))
    sec = _three_layer_section(faces[-1])
    sigma = np.empty((nx, ny, nz))
    for k in range(nz):
        sigma[:, :, k] = sec.sigma[sec.background_layer(depth[k], depth[k] + thick[k])]
    _paint_boxes(SplitMix64(seed_for(5, model_index)), sigma, 60, 4, 64, math.log10(0.3), math.log10(3000.0))
    freqs = [10.0 ** (-4 + 6 * k / 11) for k in range(12)]
    return CartesianModel(nx, ny, 100.0, 100.0, depth, thick, sigma, sec, "C5", freqs)


def small_model(nx=4, ny=4, nz=2, dx=200.0, dz=100.0, contrast=10.0, seed=1) -> CartesianModel:
    """A tiny layered model for unit tests: uniform dz, background [air | 100 ohm-m slab | half-space],
    a random conductivity perturbation of the given contrast."""
    rng = SplitMix64(seed)
    depth = np.arange(nz) * dz
    thick = np.full(nz, dz)
    sec = Section1D(0.0, [0.0, nz * dz, 0.0], [0.0, 1e-2, 1e-3])
    sigma = np.empty((nx, ny, nz))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                sigma[i, j, k] = 1e-2 * contrast ** (rng.uniform() * 2 - 1)
    return CartesianModel(nx, ny, dx, dx, depth, thick, sigma, sec, "small", [1.0])
