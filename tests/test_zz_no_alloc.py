"""The allocation-free level-field entry point (xtr_op_level_field_noalloc_device) against the
allocating one (xtr_op_level_field_device), through the C ABI: two calls in a row on different
grids, the second smaller, each equal to the allocating entry point's field; and the refusal of a
grid whose scratch exceeds the 128 MiB array."""
import ctypes as C

import numpy as np
import pytest

from extreme_amd import cartesian_model as cm
from extreme_amd import forward as F
from extreme_amd import xtreme as X

pytestmark = pytest.mark.gpu


def _fields(p, entry):
    torch = p.torch
    out = []
    for pol in (0, 1):
        js = torch.from_numpy(p.js[pol]).to("cuda")
        jq = p.op.jq(p.chi[pol], js)
        for electric in (1, 0):
            field = torch.full((p.model.nx, p.model.ny, 3), float("nan"), dtype=torch.complex128, device="cuda")
            rc = entry(p.op._h, electric, C.c_void_p(jq.data_ptr()), C.c_void_p(field.data_ptr()))
            X._check(rc)
            out.append(field.cpu().numpy())
    return out


def test_noalloc_level_field_matches_allocating(torch_cuda):
    L = X.lib()
    runs = [(cm.config_c1(), 0.1), (cm.small_model(8, 8, 4, contrast=30.0, seed=7), 1.0)]
    for k, (m, f) in enumerate(runs):          # the second call is on a smaller grid
        p = F.FrequencySolve(m, f)
        p.solve()
        p.build_level(0.0)
        got = _fields(p, L.xtr_op_level_field_noalloc_device)
        ref = _fields(p, L.xtr_op_level_field_device)
        for a, b in zip(got, ref):
            assert np.isfinite(a).all()
            assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))
        print(f"call {k}: {m.nx}x{m.ny}x{m.nz} fields equal the allocating entry point's", flush=True)


def test_noalloc_level_field_refuses_large_grid(torch_cuda):
    torch = torch_cuda
    m = cm.small_model(128, 128, 48, seed=3)     # 4 Nx Ny * 3 (Nz + 1) = 9.6 Mi elements > 8 Mi
    op = X.ConvolutionOperator(m, cm.omega_of(1.0), max_rhs=1)
    jq = torch.zeros(m.n, dtype=torch.complex128, device="cuda")
    field = torch.zeros((m.nx, m.ny, 3), dtype=torch.complex128, device="cuda")
    rc = X.lib().xtr_op_level_field_noalloc_device(op._h, 1, C.c_void_p(jq.data_ptr()), C.c_void_p(field.data_ptr()))
    assert rc != 0
    buf = C.create_string_buffer(512)
    X.lib().xtr_last_error(buf, 512)
    assert b"128 MiB" in buf.value
    assert torch.count_nonzero(field).item() == 0
