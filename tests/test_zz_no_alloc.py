"""fh_norm_factor_device_noalloc (no counterpart in the reference; for
callers that cannot allocate): through the C-ABI, its result on device
arrays equals fh_norm_factor's on the same inputs (the same bits: the same
per-block partials and finalize), on two calls in a row, the second on a
smaller mesh, and on a coupled mesh; a mesh above the scratch maximum is
refused.  The scratch array keeps the previous call's contents, so the
second, smaller call also checks that the state is cleared."""
import ctypes

import numpy as np
import pytest

import foamhip as fh
from foamhip.cases import BoxCase
from meshes import cyclic_system, random_graph, spd_coeffs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def noalloc(m, psi, src):
    d_psi = torch.from_numpy(np.ascontiguousarray(psi)).to("cuda")
    d_src = torch.from_numpy(np.ascontiguousarray(src)).to("cuda")
    d_out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_norm_factor_device_noalloc(
        m.handle, ctypes.c_void_p(d_psi.data_ptr()), ctypes.c_void_p(d_src.data_ptr()),
        ctypes.c_void_p(d_out.data_ptr())))
    torch.cuda.synchronize()
    return float(d_out.cpu().numpy()[0])


def case_box(n, seed):
    c = BoxCase(n)
    gamma = c.lognormal(seed, 1.0, 4.0)
    diag, upper, source = c.laplacian_system(gamma, c.rhs(seed + 1))
    m = fh.lduMatrix(fh.lduAddressing(c.N, c.lower, c.upper), diag, upper)
    psi = np.random.default_rng(seed).standard_normal(c.N)
    return m, psi, source


def case_graph(n, seed):
    rng = np.random.default_rng(seed)
    n, lower, upper = random_graph(n, 5.0, rng, shuffle_within_owner=True)
    diag, up = spd_coeffs(n, lower, upper, rng)
    m = fh.lduMatrix(fh.lduAddressing(n, lower, upper), diag, up)
    return m, rng.standard_normal(n), rng.standard_normal(n)


def test_noalloc_norm_factor_matches():
    for m, psi, src in [case_box(48, 3), case_graph(3000, 4), case_box(20, 5)]:
        ref = m.normFactor(psi, src)
        got = noalloc(m, psi, src)
        assert got == ref, (got, ref)


def test_noalloc_norm_factor_coupled():
    n, lower, upper, diag, up, lo, ifc, _, _ = cyclic_system("sym", 3, (24, 16, 12),
                                                           with_dense=False)
    mesh = fh.lduAddressing(n, lower, upper, interfaces=ifc)
    m = fh.lduMatrix(mesh, diag, up)
    m.set_interfaces(ifc["bouCoeffs"])
    rng = np.random.default_rng(6)
    psi, src = rng.standard_normal(n), rng.standard_normal(n)
    assert noalloc(m, psi, src) == m.normFactor(psi, src)


def test_noalloc_norm_factor_refuses_above_maximum():
    n = (1 << 24) + 1
    e = np.zeros(0, np.int32)
    m = fh.lduMatrix(fh.lduAddressing(n, e, e), np.ones(n))
    d = torch.zeros(n, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = fh.lib().fh_norm_factor_device_noalloc(m.handle, ctypes.c_void_p(d.data_ptr()),
                                                ctypes.c_void_p(d.data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()))
    assert st == 1
    assert "exceed" in fh.lib().fh_last_error().decode()
