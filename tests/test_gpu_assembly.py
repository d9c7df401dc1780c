"""GPU parity of the finite-volume assembly kernels against the oracle,
through the C-ABI: fvm::laplacian (gaussLaplacianScheme.C:45-89), upwind
convection minus diffusion (gaussConvectionScheme.C:119-170) and the Gauss
gradient (gaussGrad.C:30-75, gradScheme.C:123-148).  Each kernel keeps the
reference's per-cell order, so every output is compared bitwise.  Device
arrays are torch tensors; the library reads and writes them in place."""
import ctypes

import numpy as np
import pytest

import foamhip as fh
import oracle as O
from foamhip.cases import BoxCase

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(x, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(x if dtype is None else x.astype(dtype)))
    return t.to("cuda")


def P(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


SHAPES = [(9, 7, 5), (16, 16, 16), (33, 4, 3)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("add_diag", [0, 1])
def test_laplacian_assembly_bitwise(shape, add_diag):
    c = BoxCase(*shape, fixed_patches=(0, 3))
    rng = np.random.default_rng(sum(shape))
    gamma = np.exp(rng.standard_normal(c.N))
    c.bValue = rng.standard_normal(c.nB)
    a = O.Addr(c.N, c.lower, c.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, c.weights, c.magSf, c.deltaCoeffs,
                                               c.bFaceCells, c.bMagSf, c.bDeltaCoeffs, c.bType,
                                               c.bValue)
    rhs = rng.standard_normal(c.N)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    bnd = fh.Boundary(mesh, c.bFaceCells)
    m = fh.lduMatrix(mesh)
    d_ic = torch.empty(c.nB, dtype=torch.float64, device="cuda")
    d_bc = torch.empty_like(d_ic)
    d_src = dev(rhs)
    # inputs held by name until the call returns
    ins = [dev(gamma), dev(c.weights), dev(c.magSf), dev(c.deltaCoeffs), dev(c.bMagSf),
           dev(c.bDeltaCoeffs), dev(c.bType, np.int32), dev(c.bValue)]
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_assemble_laplacian(m.handle, bnd.handle, *[P(t) for t in ins],
                                             P(d_ic), P(d_bc), P(d_src), add_diag))
    np.testing.assert_array_equal(host(d_ic), ic[:c.nB])
    np.testing.assert_array_equal(host(d_bc), bc[:c.nB])
    d, u, lo = m.coeffs()
    np.testing.assert_array_equal(u, upper[:c.F])
    np.testing.assert_array_equal(lo, upper[:c.F])
    d2, s2 = O.add_boundary(c.bFaceCells, ic[:c.nB], bc[:c.nB], diag, rhs)
    np.testing.assert_array_equal(d, d2 if add_diag else diag)
    # the source receives the boundary coefficients in either case
    np.testing.assert_array_equal(host(d_src), s2)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("D", [1e-3, 0.5])
def test_convection_diffusion_assembly_bitwise(shape, D):
    c = BoxCase(*shape, fixed_patches=(0, 1))
    rng = np.random.default_rng(7 + sum(shape))
    c.bValue = rng.standard_normal(c.nB)
    phi, bPhi = c.fourier_flux(3)
    a = O.Addr(c.N, c.lower, c.upper)
    diag, upper, lower, ic, bc = O.convection_diffusion(a, c, phi, bPhi, D)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    bnd = fh.Boundary(mesh, c.bFaceCells)
    m = fh.lduMatrix(mesh)
    d_ic = torch.empty(c.nB, dtype=torch.float64, device="cuda")
    d_bc = torch.empty_like(d_ic)
    d_phi = dev(phi)
    ins = [dev(c.magSf), dev(c.deltaCoeffs), dev(bPhi), dev(c.bMagSf), dev(c.bDeltaCoeffs),
           dev(c.bType, np.int32), dev(c.bValue)]
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_assemble_convection_diffusion(
        m.handle, bnd.handle, P(d_phi), ctypes.c_double(D), *[P(t) for t in ins], P(d_ic),
        P(d_bc), None, 0))
    d, u, lo = m.coeffs()
    np.testing.assert_array_equal(d, diag)
    np.testing.assert_array_equal(u, upper[:c.F])
    np.testing.assert_array_equal(lo, lower[:c.F])
    np.testing.assert_array_equal(host(d_ic), ic[:c.nB])
    np.testing.assert_array_equal(host(d_bc), bc[:c.nB])


@pytest.mark.parametrize("shape", SHAPES)
def test_gauss_grad_bitwise(shape):
    c = BoxCase(*shape)
    rng = np.random.default_rng(11 + sum(shape))
    phi = rng.standard_normal(c.N)
    bPhi = rng.standard_normal(c.nB)
    bSnGrad = rng.standard_normal(c.nB)
    a = O.Addr(c.N, c.lower, c.upper)
    grad, bGrad = O.grad_gauss_linear(a, phi, c.weights, c.Sf, c.V, c.bFaceCells, c.bSf,
                                      c.bMagSf, bPhi, bSnGrad)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    bnd = fh.Boundary(mesh, c.bFaceCells)
    d_grad = torch.empty(3 * c.N, dtype=torch.float64, device="cuda")
    d_bGrad = torch.empty(3 * c.nB, dtype=torch.float64, device="cuda")
    ins = [dev(phi), dev(c.weights), dev(c.Sf), dev(c.V), dev(c.bSf), dev(c.bMagSf),
           dev(bPhi), dev(bSnGrad)]
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_grad_gauss_linear(mesh.handle, bnd.handle, *[P(t) for t in ins],
                                            P(d_grad), P(d_bGrad)))
    np.testing.assert_array_equal(host(d_grad), grad)
    np.testing.assert_array_equal(host(d_bGrad), bGrad)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind,ncmpt", [("laplacian", 1), ("convdiff", 1), ("convdiff", 3)])
def test_relax_bitwise(shape, kind, ncmpt):
    """fvMatrix::relax (fvMatrix.C:594-740) on the device: the relaxed
    diagonal (returned and kept by the matrix) and the source."""
    c = BoxCase(*shape, fixed_patches=(0, 5))
    rng = np.random.default_rng(31 + sum(shape) + ncmpt)
    a = O.Addr(c.N, c.lower, c.upper)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    bnd = fh.Boundary(mesh, c.bFaceCells)
    if kind == "laplacian":
        gamma = np.exp(rng.standard_normal(c.N))
        upper, diag, ic, _ = O.laplacian_assemble(a, gamma, c.weights, c.magSf, c.deltaCoeffs,
                                                  c.bFaceCells, c.bMagSf, c.bDeltaCoeffs,
                                                  c.bType, c.bValue)
        om = O.Matrix(a, diag, upper[:c.F])
        m = fh.lduMatrix(mesh, diag, upper[:c.F])
        ic = ic[:c.nB]
    else:
        phi, bPhi = c.fourier_flux(5)
        diag, upper, lower, ic, _ = O.convection_diffusion(a, c, phi, bPhi, 0.05)
        om = O.Matrix(a, diag, upper[:c.F], lower[:c.F])
        m = fh.lduMatrix(mesh, diag, upper[:c.F], lower[:c.F])
        ic = ic[:c.nB]
    if ncmpt > 1:
        ic = np.stack([ic, ic * rng.uniform(0.5, 1.5, c.nB), -np.abs(ic)], axis=1)
    shape_f = (c.N,) if ncmpt == 1 else (c.N, ncmpt)
    psi = rng.standard_normal(shape_f)
    src = rng.standard_normal(shape_f)
    d_ref, s_ref = om.relax(c.bFaceCells, ic, 0.7, psi, src)
    d_ic, d_psi, d_src = dev(ic), dev(psi), dev(src)
    d_diag = torch.empty(c.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_relax(m.handle, bnd.handle, ctypes.c_double(0.7), ncmpt, P(d_ic),
                                P(d_psi), P(d_src), P(d_diag)))
    np.testing.assert_array_equal(host(d_diag), d_ref)
    np.testing.assert_array_equal(host(d_src), s_ref)
    np.testing.assert_array_equal(m.coeffs()[0], d_ref)
    # the relaxed matrix solves like one set from the host with that diagonal
    x = rng.standard_normal(c.N)
    np.testing.assert_array_equal(m.Amul(x), O.Matrix(a, d_ref, upper[:c.F],
                                                      None if kind == "laplacian"
                                                      else lower[:c.F]).Amul(x))
    # alpha <= 0 leaves everything as it is
    before = host(d_src).copy()
    fh._check(fh.lib().fh_relax(m.handle, bnd.handle, ctypes.c_double(0.0), ncmpt, P(d_ic),
                                P(d_psi), P(d_src), P(d_diag)))
    np.testing.assert_array_equal(host(d_src), before)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ncmpt,with_corr", [(1, False), (1, True), (3, True)])
def test_flux_bitwise(shape, ncmpt, with_corr):
    """fvMatrix::flux (fvMatrix.C:953-1028) of an assembled asymmetric
    matrix: internal and boundary face fluxes."""
    c = BoxCase(*shape, fixed_patches=(0, 2))
    rng = np.random.default_rng(41 + sum(shape) + ncmpt)
    c.bValue = rng.standard_normal(c.nB)
    a = O.Addr(c.N, c.lower, c.upper)
    phi, bPhi = c.fourier_flux(9)
    diag, upper, lower, ic, bc = O.convection_diffusion(a, c, phi, bPhi, 0.02)
    ic, bc = ic[:c.nB], bc[:c.nB]
    if ncmpt > 1:
        ic = np.stack([ic, 0.5 * ic, -ic], 1)
        bc = np.stack([bc, -bc, 2 * bc], 1)
    om = O.Matrix(a, diag, upper[:c.F], lower[:c.F])
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    bnd = fh.Boundary(mesh, c.bFaceCells)
    m = fh.lduMatrix(mesh, diag, upper[:c.F], lower[:c.F])
    shp = lambda n: (n,) if ncmpt == 1 else (n, ncmpt)
    psi = rng.standard_normal(shp(c.N))
    corr = rng.standard_normal(shp(c.F)) if with_corr else None
    bCorr = rng.standard_normal(shp(c.nB)) if with_corr else None
    fl, bf = om.flux(c.bFaceCells, ic, bc, psi, corr, bCorr)
    d_psi, d_ic, d_bc = dev(psi), dev(ic), dev(bc)
    d_corr = dev(corr) if with_corr else None
    d_bCorr = dev(bCorr) if with_corr else None
    d_fl = torch.empty(c.F * ncmpt, dtype=torch.float64, device="cuda")
    d_bf = torch.empty(c.nB * ncmpt, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_flux(m.handle, bnd.handle, ncmpt, P(d_psi), P(d_ic), P(d_bc),
                               P(d_corr), P(d_bCorr), P(d_fl), P(d_bf)))
    np.testing.assert_array_equal(host(d_fl), fl.reshape(-1))
    np.testing.assert_array_equal(host(d_bf), bf.reshape(-1))


@pytest.mark.parametrize("case", ["box", "rand_shuffled"])
def test_h_device_matches_host(case):
    """fh_h_device / fh_face_h_device on device arrays give the host entry
    points' results (which match the oracle bitwise), also where the library
    numbering differs from the caller's."""
    from meshes import asym_coeffs, box, random_graph
    rng = np.random.default_rng(51)
    n, lo, up = box(13, 9, 7) if case == "box" else random_graph(3000, 6, rng,
                                                                 shuffle_within_owner=True)
    d, u, l = asym_coeffs(n, lo, up, rng)
    mesh = fh.lduAddressing(n, lo, up)
    m = fh.lduMatrix(mesh, d, u, l)
    for nc in (1, 3):
        X = rng.standard_normal((n, nc) if nc > 1 else n)
        d_x = dev(X)
        d_h = torch.empty(n * nc, dtype=torch.float64, device="cuda")
        d_f = torch.empty(len(lo) * nc, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        fh._check(fh.lib().fh_h_device(m.handle, nc, P(d_x), P(d_h)))
        fh._check(fh.lib().fh_face_h_device(m.handle, nc, P(d_x), P(d_f)))
        np.testing.assert_array_equal(host(d_h), m.H(X).reshape(-1))
        np.testing.assert_array_equal(host(d_f), m.faceH(X).reshape(-1))
