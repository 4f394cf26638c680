"""The reference's on-disk formats: `.xproj` projects, `.xmodel` v2.0 models with plain-text or
applique anomaly layers, and the MT result files `fields_freq*` / `responses_freq*`.

A project written for the reference runs here unchanged (`run_project`), and the result files
have the reference's column layout and number formats, so the two engines' outputs can be compared
file against file. Readers follow `Extreme.Cartesian/Project/ProjectReader.cs:27-128`,
`Forward/Project/ForwardSettingsReader.cs:10-20`, `Model/LoadAndSave/ModelReader.cs:21-172` and
`AnomalyLoaderUtils.cs:14-171`; writers follow
`Extreme.Cartesian.Magnetotellurics/Results/PlainTextExporter.cs:28-448` and
`AllFieldsUtils.cs:11-37`.
"""
from __future__ import annotations

import math
import os
import re
import time
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import cartesian_model as cm
from .xtreme import ForwardSettings


# ---------------------------------------------------------------------------- project (.xproj)

@dataclass
class ObservationLevel:
    """ObservationLevel: a plane of sites at depth z under the anomaly's cell centres, shifted
    laterally by (x_shift, y_shift) metres."""
    x_shift: float
    y_shift: float
    z: float
    name: str = ""


@dataclass
class ObservationSite:
    x: float
    y: float
    z: float
    name: str = ""


@dataclass
class SourceLayer:
    x_shift: int
    y_shift: int
    z: float


@dataclass
class ExtremeProject:
    """ExtremeProject as ProjectReader builds it (frequencies, observations, sources, settings)."""
    frequencies: list
    levels: list
    sites: list
    sources: list
    model_file: str | None
    results_path: str | None
    settings: ForwardSettings
    path: str = ""

    def resolve(self, relative: str) -> str:
        """A path from the project file, relative to the project's directory."""
        return os.path.join(os.path.dirname(os.path.abspath(self.path)), relative)


def _xml_root(path: str) -> ET.Element:
    with open(path, "rb") as f:
        text = f.read().decode("utf-8-sig")
    return ET.fromstring(text)


def _child(elem, name):
    return None if elem is None else elem.find(name)


def _child_number(elem, name, cast, default):
    """ElementAsIntOrNull / ElementAsDoubleOrNull with the reader's fallback."""
    c = _child(elem, name)
    if c is None or c.text is None or not c.text.strip():
        return default
    return cast(c.text.strip())


def _attr(elem, name, cast=float, default=None):
    v = elem.get(name)
    if v is None or not v.strip():
        if default is None:
            raise ValueError(f"missing attribute {name} on <{elem.tag}>")
        return default
    return cast(v.strip())


def read_forward_settings(xelem) -> ForwardSettings:
    """ForwardSettingsReader.FromXElement (Forward/Project/ForwardSettingsReader.cs:10-20): missing
    elements take the reader's defaults (10, 10, 10, 10, 1e-10), not the class's."""
    return ForwardSettings(
        residual=_child_number(xelem, "Residual", float, 1e-10),
        inner_buffer_length=_child_number(xelem, "InnerBufferLength", int, 10),
        outer_buffer_length=_child_number(xelem, "OuterBufferLength", int, 10),
        max_repeats_number=_child_number(xelem, "MaxRepeatsNumber", int, 10),
        number_of_hankels=_child_number(xelem, "NumberOfHankels", int, 10))


def read_project(path: str) -> ExtremeProject:
    """ProjectReader.FromXDocument (Project/ProjectReader.cs:27-50): frequencies, tablets
    (observation levels <T z xShift yShift name>), sites (<S x y z name>), source layers (<L z xShift
    yShift>), the Forward settings, the model file and the results path."""
    root = _xml_root(path)
    if root.tag != "ExtremeProject":
        raise ValueError("No ExtremeProject section in project file")
    xf = root.find("Frequencies")
    freqs = [float(f.text.strip()) for f in xf.findall("F")] if xf is not None else []
    obs = root.find("Observations")
    levels = [ObservationLevel(_attr(t, "xShift", float, 0.0), _attr(t, "yShift", float, 0.0), _attr(t, "z"),
                               t.get("name", "") or "")
              for t in (_child(obs, "Tablets") if _child(obs, "Tablets") is not None else [])]
    sites = [ObservationSite(_attr(s, "x"), _attr(s, "y"), _attr(s, "z"), s.get("name", "") or "")
             for s in (_child(obs, "Sites") if _child(obs, "Sites") is not None else [])]
    src = _child(root.find("Sources"), "Layers")
    sources = [SourceLayer(_attr(l, "xShift", int, 0), _attr(l, "yShift", int, 0), _attr(l, "z"))
               for l in (src if src is not None else [])]
    settings = read_forward_settings(_child(root.find("Settings"), "Forward"))

    def text(name):
        e = root.find(name)
        return e.text.strip() if e is not None and e.text is not None else None

    return ExtremeProject(freqs, levels, sites, sources, text("ModelFile"), text("ResultsPath"), settings, path)


# ---------------------------------------------------------------------------- model (.xmodel v2.0)

def _doubles(line: str) -> list:
    """AnomalyLoaderUtils.GetAllDouble: space/tab separated, Fortran 'D' exponents accepted."""
    return [float(v.replace("D", "E").replace("d", "e")) for v in re.split(r"[ \t]+", line.strip()) if v]


def read_plain_text_layer(path: str, nx: int, ny: int) -> np.ndarray:
    """AnomalyLoaderUtils.ReadAnomalyDataFromPlainText (:41-65): nx lines (i, top to bottom) of ny
    values (j, left to right)."""
    with open(path) as f:
        lines = [f.readline() for _ in range(nx)]
    out = np.empty((nx, ny))
    for i, line in enumerate(lines):
        if not line:
            raise ValueError("model with illegal Nx value")
        v = _doubles(line)
        if len(v) != ny:
            raise ValueError(f"model with illegal Ny value in line {i}")
        out[i] = v
    return out


_APPLIQUE_LINES = re.compile(r"(?P<lines>\d+)LINES:", re.IGNORECASE)
_APPLIQUE_CELLS = re.compile(r"\s*(?P<columns>\d+)\*(?P<value>[0-9]*\.?[0-9]+([eE][-+]?[0-9]+)*)", re.IGNORECASE)


def parse_applique(template: str, sigma: np.ndarray, k: int) -> None:
    """AnomalyLoaderUtils.ParseApplique (:76-171): each non-empty line `NLINES: c1*v1 c2*v2 ...`
    fills the next N values of the second index with runs of c cells of value v along the first
    index (sigma[j, i, k] = v, the reference's order), values parsed as float32."""
    ny, nx = sigma.shape[0], sigma.shape[1]
    prev = 0
    for line in (s.strip() for s in template.split("\n")):
        if not line:
            continue
        m = _APPLIQUE_LINES.search(line)
        if m is None:
            raise ValueError(f"applique line without NLINES: {line!r}")
        lines = int(m.group("lines"))
        runs = [(int(c.group("columns")), float(np.float32(c.group("value"))))
                for c in _APPLIQUE_CELLS.finditer(line[m.end():])]
        if prev + lines > nx:
            raise ValueError("Applique nx is out of range")
        if sum(c for c, _ in runs) > ny:
            raise ValueError("Applique ny is out of range")
        for i in range(prev, prev + lines):
            col = 0
            for c, v in runs:
                sigma[col:col + c, i, k] = v
                col += c
        prev += lines


def read_model(path: str, frequencies=None, name: str | None = None) -> cm.CartesianModel:
    """ModelReader.CartesianModelFromX + PopulateAnomaly (Model/LoadAndSave/ModelReader.cs:21-172):
    version 2.0 only; lateral dimensions; the background section (zeroLevel, layers of thickness
    and sigma; thickness 0 marks the half-spaces); anomaly layers (depth, thickness) with values
    from a plain-text file next to the model and/or an applique."""
    root = _xml_root(path)
    if root.tag != "CartesianModel":
        raise ValueError("not a CartesianModel file")
    if root.get("version") != "2.0":
        raise ValueError("Wrong version")
    ld = root.find("LateralDimensions")
    if ld is None:
        raise ValueError("missing lateral dimensions")
    nx, ny = _attr(ld, "nx", int), _attr(ld, "ny", int)
    dx, dy = _attr(ld, "dx"), _attr(ld, "dy")
    bg = root.find("Background1DSection")
    if bg is None:
        raise ValueError("background layers are not presented")
    zero = _attr(bg, "zeroLevel", float, 0.0)
    layers = bg.findall("Layer")
    section = cm.Section1D(zero, [_attr(l, "thickness") for l in layers], [_attr(l, "sigma") for l in layers])
    xan = root.find("Anomaly")
    xlayers = xan.findall("AnomalyLayer") if xan is not None else []
    depth = np.array([_attr(l, "depth") for l in xlayers], dtype=float)
    thick = np.array([_attr(l, "thickness") for l in xlayers], dtype=float)
    sigma = np.zeros((nx, ny, len(xlayers)))
    base = os.path.dirname(os.path.abspath(path))
    for k, xl in enumerate(xlayers):
        ff = xl.find("FromFile")
        if ff is not None and ff.get("fileType") == "plain-text" and ff.get("fileName"):
            sigma[:, :, k] = read_plain_text_layer(os.path.join(base, ff.get("fileName")), nx, ny)
        ap = xl.find("Applique")
        if ap is not None:
            parse_applique(ap.text or "", sigma, k)
    return cm.CartesianModel(nx, ny, dx, dy, depth, thick, sigma, section,
                             name if name is not None else os.path.splitext(os.path.basename(path))[0],
                             list(frequencies or []))


def _num(v: float) -> str:
    """A double as .NET's default ToString() writes it: the shortest round-trip form."""
    r = repr(float(v))
    return r[:-2] if r.endswith(".0") else r


def write_plain_text_layer(sigma: np.ndarray, k: int, path: str) -> None:
    """AnomalyLoaderUtils.WriteAnomalyDataToPlainText (:14-26): one line per i, `value ` per j."""
    with open(path, "w") as f:
        for i in range(sigma.shape[0]):
            f.write("".join(f"{_num(v)} " for v in sigma[i, :, k]) + "\n")


def write_model(model: cm.CartesianModel, path: str, layer_pattern: str = "anomaly_layer_{:04d}.dat") -> None:
    """A v2.0 model file with one plain-text layer file per anomaly layer beside it (the layout of
    the reference's example projects)."""
    base = os.path.dirname(os.path.abspath(path))
    lines = ['<?xml version="1.0" encoding="utf-8"?>', '<CartesianModel version="2.0">',
             f'  <LateralDimensions nx="{model.nx}" ny="{model.ny}" dx="{_num(model.dx)}" dy="{_num(model.dy)}" />',
             f'  <Background1DSection zeroLevel="{_num(model.section.zero_level)}">']
    for t, s in zip(model.section.thickness, model.section.sigma):
        lines.append(f'    <Layer thickness="{_num(t)}" sigma="{_num(s)}" />')
    lines += ["  </Background1DSection>", "  <Anomaly>"]
    for k in range(model.nz):
        fname = layer_pattern.format(k)
        write_plain_text_layer(model.sigma, k, os.path.join(base, fname))
        lines += [f'    <AnomalyLayer depth="{_num(model.depth[k])}" thickness="{_num(model.thick[k])}">',
                  f'      <FromFile fileType="plain-text" fileName="{fname}" />', "    </AnomalyLayer>"]
    lines += ["  </Anomaly>", "</CartesianModel>"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------------------- results (PlainTextExporter)

def net_e(v: float) -> str:
    """.NET's ToString("E"): 6 decimals and a signed exponent of at least 3 digits."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    mant, exp = f"{v:.6E}".split("E")
    return f"{mant}E{exp[0]}{exp[1:].zfill(3)}"


def frequency_to_string(frequency: float) -> str:
    """PlainTextExporter.FrequencyToString (:28-29): format 0000000000.0000000000."""
    return f"{frequency:021.10f}"


def responses_file_name(frequency: float) -> str:
    return f"responses_freq{frequency_to_string(frequency)}"


def fields_file_name(frequency: float) -> str:
    return f"fields_freq{frequency_to_string(frequency)}"


@dataclass
class LevelFields:
    """AllFieldsAtLevel for one level: normal and anomalous E, H of both polarisations on the
    level's Nx x Ny sites, each [nx, ny, 3] complex."""
    level: ObservationLevel
    normal_e: tuple
    normal_h: tuple
    anomaly_e: tuple
    anomaly_h: tuple
    sites_x: np.ndarray = field(default=None)
    sites_y: np.ndarray = field(default=None)


def site_coordinates(model: cm.CartesianModel, level: ObservationLevel):
    """AllFieldsUtils.GetSite (:11-37): x = dx i + dx/2 + xShift, y = dy j + dy/2 + yShift."""
    x = model.dx * np.arange(model.nx) + model.dx / 2 + level.x_shift
    y = model.dy * np.arange(model.ny) + model.dy / 2 + level.y_shift
    return x, y


def _head(f) -> None:
    f.write("%" + "{0} {1} {2} {3} ".format("x".rjust(10), "y".rjust(10), "z".rjust(10), "freq".rjust(14)))


def _pad(f, text: str) -> None:
    f.write(f"{text.rjust(14)} ")


def _coords(f, x, y, z, frequency) -> None:
    """WriteFreqAndCoord (:325-338): " x y z " with F2 values padded to 10, then the frequency."""
    f.write(" " + f"{x:.2f}".rjust(10) + " " + f"{y:.2f}".rjust(10) + " " + f"{z:.2f}".rjust(10) + " ")
    f.write(f"{net_e(frequency).rjust(14)} ")


def _cplx(f, c) -> None:
    f.write(f"{net_e(c.real).rjust(14)} {net_e(c.imag).rjust(14)} ")


def impedance_tensor(e1, e2, h1, h2):
    """ResponseFunctionsCalculator.CalculateTensor (Magnetotellurics/ResponseFunctionsCalculator.cs:36-51)."""
    det = h1[..., 0] * h2[..., 1] - h2[..., 0] * h1[..., 1]
    hi11, hi12 = h2[..., 1] / det, -h2[..., 0] / det
    hi21, hi22 = -h1[..., 1] / det, h1[..., 0] / det
    return (e1[..., 0] * hi11 + e2[..., 0] * hi21, e1[..., 0] * hi12 + e2[..., 0] * hi22,
            e1[..., 1] * hi11 + e2[..., 1] * hi21, e1[..., 1] * hi12 + e2[..., 1] * hi22)


def tipper(h1, h2):
    """ResponseFunctionsCalculator.CalculateTipper (:97-110): [Wzx Wzy] = [h1z h2z][H1 H2]^-1."""
    det = h1[..., 0] * h2[..., 1] - h2[..., 0] * h1[..., 1]
    hi11, hi12 = h2[..., 1] / det, -h2[..., 0] / det
    hi21, hi22 = -h1[..., 1] / det, h1[..., 0] / det
    return h1[..., 2] * hi11 + h2[..., 2] * hi21, h1[..., 2] * hi12 + h2[..., 2] * hi22


_VEC_NAMES = [(f"{kind}_pol{pol}_{fld}{c}") for kind, pol in (("N", "X"), ("N", "Y"), ("A", "X"), ("A", "Y"))
              for fld in ("E", "H") for c in ("x", "y", "z")]


def export_raw_fields(path: str, model: cm.CartesianModel, levels: list, frequency: float) -> None:
    """PlainTextExporter.ExportRawFields (:136-235): per site normal E1 H1 E2 H2, then anomalous
    E1 H1 E2 H2, each vector x, y, z as re, im."""
    with open(path, "w") as f:
        _head(f)
        for name in _VEC_NAMES:
            _pad(f, name + "_re")
            _pad(f, name + "_im")
        f.write("\n")
        for lf in levels:
            xs, ys = site_coordinates(model, lf.level)
            vecs = [lf.normal_e[0], lf.normal_h[0], lf.normal_e[1], lf.normal_h[1],
                    lf.anomaly_e[0], lf.anomaly_h[0], lf.anomaly_e[1], lf.anomaly_h[1]]
            for i in range(model.nx):
                for j in range(model.ny):
                    _coords(f, xs[i], ys[j], lf.level.z, frequency)
                    for v in vecs:
                        for c in range(3):
                            _cplx(f, v[i, j, c])
                    f.write("\n")


def export_mt_responses(path: str, model: cm.CartesianModel, levels: list, frequency: float) -> None:
    """PlainTextExporter.ExportMtResponses (:238-320): Z, rho_a, phase (radians), magnetic tipper,
    electric tipper from the full fields (normal + anomalous) at each site."""
    omega = cm.omega_of(frequency)
    with open(path, "w") as f:
        _head(f)
        for n in ("Z_XX", "Z_XY", "Z_YX", "Z_YY"):
            _pad(f, n + "_re")
            _pad(f, n + "_im")
        for n in ("Rho_XX", "Rho_XY", "Rho_YX", "Rho_YY", "Phs_XX", "Phs_XY", "Phs_YX", "Phs_YY"):
            _pad(f, n)
        for n in ("W_ZX", "W_ZY", "EW_ZX", "EW_ZY"):
            _pad(f, n + "_re")
            _pad(f, n + "_im")
        f.write("\n")
        for lf in levels:
            xs, ys = site_coordinates(model, lf.level)
            e1, e2 = lf.normal_e[0] + lf.anomaly_e[0], lf.normal_e[1] + lf.anomaly_e[1]
            h1, h2 = lf.normal_h[0] + lf.anomaly_h[0], lf.normal_h[1] + lf.anomaly_h[1]
            z = impedance_tensor(e1, e2, h1, h2)
            w = tipper(h1, h2)
            ew = tipper(e1, e2)
            for i in range(model.nx):
                for j in range(model.ny):
                    _coords(f, xs[i], ys[j], lf.level.z, frequency)
                    for c in z:
                        _cplx(f, c[i, j])
                    for c in z:
                        f.write(f"{net_e(abs(c[i, j]) ** 2 / (omega * cm.MU0)).rjust(14)} ")
                    for c in z:
                        f.write(f"{net_e(np.angle(c[i, j])).rjust(14)} ")
                    for c in (*w, *ew):
                        _cplx(f, c[i, j])
                    f.write("\n")


def read_result_table(path: str):
    """Reads a fields or responses file back: (column names, rows as float array)."""
    with open(path) as f:
        head = f.readline()
        rows = [[float(v) for v in line.split()] for line in f if line.strip()]
    return head.lstrip("%").split(), np.array(rows)


# ---------------------------------------------------------------------------- driver

COMMON_COLUMNS = (("Date", 15), ("Time", 12), ("nx", 6), ("ny", 6), ("nz", 6), ("nhank", 6), ("mpi_size", 9),
                  ("nthreads", 9), ("AtoA_Green", 15), ("polX", 15), ("polY", 15), ("AtoO_Green", 15),
                  ("AtoO_calc_X", 15), ("AtoO_calc_Y", 15))


def common_dat_line(model: cm.CartesianModel, nhank: int, atoa_s: float, pol_s, atoo_s: float, atoo_calc_s,
                    when=None, mpi_size: int = 1, nthreads: int = 1) -> str:
    """One row of common.dat, as ProfilerResultsTextExporter.SaveProfilingResultsToCommon writes it
    (EndUser/Profiling/ProfilerResultsTextExporter.cs:266-305): short date and time of the invariant
    culture (MM/dd/yyyy, HH:mm), the grid, NumberOfHankels, MPI size and threads, then seconds as
    "0.00000" for the AtoA Green tensor, SolveCie of each polarisation, the AtoO Green tensors and the
    AtoO fields of each polarisation; every field left-aligned in its column width."""
    import datetime
    now = when or datetime.datetime.now()
    vals = [now.strftime("%m/%d/%Y"), now.strftime("%H:%M"), str(model.nx), str(model.ny), str(model.nz),
            str(nhank), str(mpi_size), str(nthreads)]
    vals += [f"{t:.5f}" for t in (atoa_s, pol_s[0], pol_s[1], atoo_s, atoo_calc_s[0], atoo_calc_s[1])]
    return "".join(v.ljust(w) for v, (_, w) in zip(vals, COMMON_COLUMNS)) + "\n"


def append_common_dat(path: str, line: str) -> None:
    """Appends a row, writing the header first when the file does not exist yet."""
    if not os.path.exists(path):
        with open(path, "a", newline="") as f:
            f.write("".join(name.ljust(w) for name, w in COMMON_COLUMNS) + "\n")
    with open(path, "a", newline="") as f:
        f.write(line)


# Events of the per-frequency profiling report, in the order and nesting of
# ProfilerResultsTextExporter.SaveProfilingResultsTo (EndUser/Profiling/ProfilerResultsTextExporter.cs:37-55,
# 160-238): top-level events with their sub-events; an event without a measurement is left out, as the
# exporter's WriteIfExist leaves out events the profiler never recorded.
PROFILING_SECTIONS = (
    ("CustomFft", ("CustomFftInitialTranspose", "CustomFftFourierY", "CustomFftBlockTransposeYtoX",
                   "CustomFftDistributedTranspose", "CustomFftFourierX", "CustomFftBlockTransposeXtoY",
                   "CustomFftFinalTranspose")),
    ("GreenAtoATotal", (("GreenScalarAtoA", ("GreenScalarAtoACalcCalc", "GreenScalarAtoAUnion",
                                             "GreenScalarAtoASegments",
                                             ("GreenScalarAtoACommunicate", ("GreenScalarAtoATrans",)))),
                        ("GreenTensorAtoA", ("GreenTensorAtoACalc", "GreenTensorAtoAFft")))),
    ("SolveCie", ("CalcDotProduct",
                  ("ApplyOperatorA", ("OperatorAApplyR", "OperatorAPrepareForForwardFft", "OperatorAForwardFft",
                                      "OperatorAMultiplication", "OperatorABackwardFft",
                                      "OperatorAExtractAfterBackwardFft", "OperatorAFinish")))),
    ("ObservationsFullCalculation", ()),
    ("CalcChi0", ()),
    ("CalcJScattered", ()),
    ("CalcEScattered", ()),
    ("FftwPlanCalculation", ()),
)


def profiling_file_name(frequency: float, nmpi: int = 1, nthreads: int = 1) -> str:
    """profiling_mpi{nMpi:0000}_threads{nThreads}_freq{frequency:####0.0000}.txt
    (EndUser/ExtremeMt/ExtremeMtSolverRunner.cs:118,135)."""
    return f"profiling_mpi{nmpi:04d}_threads{nthreads}_freq{frequency:.4f}.txt"


def profiling_report(model: cm.CartesianModel, stats: dict, nmpi: int = 1, nthreads: int = 1, when=None) -> str:
    """Text of the per-frequency profiling file, laid out as ProfilerResultsTextExporter.
    SaveProfilingResultsTo writes it (ProfilerResultsTextExporter.cs:37-55, 57-83, 125-151, 213-264):
    the model block, the counts of operator multiplications and dot products, the total SOLVING time,
    then each top-level event with its time, count and percent of ForwardSolving and its sub-events'
    time, percent of their parent and count, with "Covered" sums, and the total covered percent.
    stats maps an event name to (seconds, count); ForwardSolving must be present."""
    import datetime
    now = when or datetime.datetime.now()
    total = stats["ForwardSolving"][0]
    out = []
    w = out.append
    w(f"DateTime: {now.strftime('%m/%d/%Y %H:%M:%S')}\n")
    w("model parameters:\n")
    w(f"nx = {model.nx}, ny = {model.ny}, nz = {model.nz}\n")
    w(f"Number of MPI processes = {nmpi}\n")
    w(f"Number of threads = {nthreads}\n")
    w("\n")
    if "OperatorAMultiplication" in stats:
        w(f"Total number of multiplications: {stats['OperatorAMultiplication'][1]}\n")
    if "CalcDotProduct" in stats:
        w(f"Total number of dot products: {stats['CalcDotProduct'][1]}\n")
    w("\n")
    w(f"Total SOLVING time {total:.5f}, {stats['ForwardSolving'][1]} times\n")

    def sub(parent, node, depth):
        name, children = (node, ()) if isinstance(node, str) else node
        if name not in stats:
            return 0.0
        t, n = stats[name]
        pct = 100.0 * t / stats[parent][0]
        w("\t" * depth + f"{name.ljust(30)} {t:.5f}, {pct:.1f} %, {n} times\n")
        if children:
            w("\n")
            covered = sum(sub(name, c, depth + 1) for c in children)
            if covered != 0:
                w("\t" * depth + f"\t\t\t\t\t\t\tCovered: {covered:.1f}%\n")
            w("\n")
        return pct

    percent = 0.0
    for top, children in PROFILING_SECTIONS:
        if top not in stats:
            continue
        t, n = stats[top]
        pct = 100.0 * t / total
        w("\n\n")
        w(f"{top}:\n")
        w(f"\tTotal time: {t:.4f}, {n} times, \t\t\t\t\tpercent of SOLVING: {pct:.1f}%\n")
        w("\n")
        covered = sum(sub(top, c, 1) for c in children)
        if covered != 0:
            w(f"\t\t\t\t\t\t\tCovered: {covered:.1f}%\n")
        percent += pct
    w(f"\n\n                                               Total Covered: {percent:.1f} %\n")
    return "".join(out)


def run_project(path: str, precision: str = "fp64", results_path: str | None = None, log=print) -> list:
    """ExtremeMtSolverRunner.Run (EndUser/ExtremeMt/ExtremeMtSolverRunner.cs:47-76) on the GPU:
    for each frequency, solve both polarisations, compute every observation level's fields and
    write fields_freq* and responses_freq* into the results path, the frequency's event breakdown
    into profiling_mpi0001_threads1_freq*.txt there, and append its timings to common.dat in the
    results path's parent (ExportProfiling, :114-141; one process and one host thread driving the
    GPU). Returns the fields and responses files written. Observation sites (single points off the
    level grid) are not built here."""
    from . import forward as F
    project = read_project(path)
    if project.model_file is None:
        raise ValueError("the project names no model file")
    if project.sites:
        raise NotImplementedError("observation sites are not supported")
    model = read_model(project.resolve(project.model_file), project.frequencies)
    out_dir = results_path or project.resolve(project.results_path or "results")
    os.makedirs(out_dir, exist_ok=True)
    written = []
    for freq in project.frequencies:
        t_freq = time.perf_counter()
        fs = F.FrequencySolve(model, freq, project.settings, precision=precision)
        fs.op.set_profiling(True)
        t0 = time.perf_counter()
        infos = fs.solve() if precision != "mixed" else fs.solve_mixed()
        solve_s = time.perf_counter() - t0
        stage_ms, applies = fs.op.stage_times()
        fs.op.set_profiling(False)
        t0 = time.perf_counter()
        log(f"frequency {freq}: {[getattr(i, 'number_of_iterations', getattr(i, 'refinements', None)) for i in infos]}")
        levels = []
        for lev in project.levels:
            ob = fs.observe(lev.z, lev.x_shift, lev.y_shift)
            shape = (model.nx, model.ny, 3)
            ne = tuple(np.broadcast_to(v, shape) for v in ob["E_normal"])
            nh = tuple(np.broadcast_to(v, shape) for v in ob["H_normal"])
            levels.append(LevelFields(lev, ne, nh, tuple(ob["E_anomalous"]), tuple(ob["H_anomalous"])))
        observe_s = time.perf_counter() - t0
        fpath = os.path.join(out_dir, fields_file_name(freq))
        rpath = os.path.join(out_dir, responses_file_name(freq))
        export_raw_fields(fpath, model, levels, freq)
        export_mt_responses(rpath, model, levels, freq)
        written += [fpath, rpath]
        # SolveCie per polarisation: each FP64 solve's own time (the two run concurrently); the
        # mixed-precision pair is timed together, so both columns carry that time
        pol_s = [i.seconds for i in infos] if precision != "mixed" else [solve_s, solve_s]
        # the frequency's event breakdown: operator stages from HIP events on the operator's stream
        # (device time of every apply of the solve, both polarisations together; R, the padding, the
        # crop and the last step are fused into the FFT passes, so they carry no separate events)
        st = {k: v * 1e-3 for k, v in stage_ms.items()}
        stats = {"ForwardSolving": (time.perf_counter() - t_freq, 1),
                 "GreenAtoATotal": (fs.green_seconds, 1),
                 "SolveCie": (solve_s, 2),          # one span for the two concurrent solves
                 "ApplyOperatorA": (sum(st.values()), applies),
                 "OperatorAForwardFft": (st["fwd_y"] + st["fwd_x"], applies),
                 "OperatorAMultiplication": (st["contract"], applies),
                 "OperatorABackwardFft": (st["inv_x"] + st["inv_y"], applies),
                 "ObservationsFullCalculation": (observe_s, 1),
                 "CalcChi0": (fs.chi0_seconds, 1),
                 "CalcJScattered": (fs.js_seconds, 1)}
        ppath = os.path.join(out_dir, profiling_file_name(freq))
        with open(ppath, "w", newline="") as f:
            f.write(profiling_report(model, stats))
        append_common_dat(os.path.join(os.path.dirname(os.path.abspath(out_dir)), "common.dat"),
                          common_dat_line(model, project.settings.number_of_hankels, fs.green_seconds, pol_s,
                                          fs.atoo_green_seconds, fs.atoo_field_seconds))
        del fs
    return written
