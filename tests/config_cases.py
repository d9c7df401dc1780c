"""The measured configurations at their own sizes, built the same way by the
GPU tests and by the fixture script (tests/golden/make_config_fixtures.py),
and a compact summary of a solution vector for comparing full-size
solutions without storing them.

Summary of a vector x (N values):
  * norm and sum;
  * K Gaussian projections g_k . x, g_k drawn from numpy's PCG64 seeded
    (PROJ_SEED, k): for a difference d = y - x, E[(g.d)^2] = |d|^2, so
    sqrt(mean_k (g_k.y - g_k.x)^2) estimates |y - x| (K = 32: within about
    30% with high probability);
  * x at M seeded indices, compared entry by entry.
"""
import numpy as np

import oracle as O
from foamhip.cases import BoxCase

PROJ_SEED = 20261203
K_PROJ = 32
M_SAMPLE = 2048


def projections(x):
    x = np.asarray(x, np.float64)
    out = np.empty(K_PROJ)
    for k in range(K_PROJ):
        g = np.random.default_rng([PROJ_SEED, k]).standard_normal(x.size)
        out[k] = g @ x
    return out


def sample_index(N):
    return np.random.default_rng([PROJ_SEED, 999]).integers(0, N, M_SAMPLE)


def summary(x):
    return {"norm": float(np.linalg.norm(x)), "sum": float(x.sum()),
            "proj": projections(x).tolist(), "sample": x[sample_index(x.size)].tolist()}


def rel_diff_estimate(y, s):
    """Estimated |y - x| / |x| for the vector x that summary s describes."""
    d = projections(y) - np.asarray(s["proj"])
    return float(np.sqrt(np.mean(d * d)) / s["norm"])


def c1_poisson(n=64, seed=1):
    """fvm::laplacian(1, p) on the unit n^3 box, x-min wall fixedValue 0,
    other patches zeroGradient, RHS U[-1,1] minus its mean (seed)."""
    c = BoxCase(n)
    a = O.Addr(c.N, c.lower, c.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, np.ones(c.N), c.weights, c.magSf,
                                               c.deltaCoeffs, c.bFaceCells, c.bMagSf,
                                               c.bDeltaCoeffs, c.bType, c.bValue)
    b = c.rhs(seed)
    diag, b = O.add_boundary(c.bFaceCells, ic, bc, diag, b)
    return c, a, diag, upper, b


C1_CONTROLS = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-6, "relTol": 0}


def c2_convection_diffusion(n=128, seed=2, D=1e-3):
    """Upwind convection of a seeded divergence-free velocity (8 Fourier
    modes, |U| ~ 1) minus the Laplacian with diffusivity D, x-min wall
    fixedValue 0, RHS U[-1,1] minus its mean."""
    c = BoxCase(n)
    a = O.Addr(c.N, c.lower, c.upper)
    phi, bPhi = c.fourier_flux(seed)
    diag, upper, lower, ic, bc = O.convection_diffusion(a, c, phi, bPhi, D)
    b = c.rhs(seed)
    diag, b = O.add_boundary(c.bFaceCells, ic, bc, diag, b)
    return c, a, diag, upper, lower, b


C2_CONTROLS = {"solver": "PBiCGStab", "preconditioner": "DILU", "tolerance": 1e-8,
               "relTol": 0}


def c2_grad_inputs(c, seed=3):
    """A seeded scalar field and its boundary values and normal gradients
    as the x-min fixedValue 0 wall and zeroGradient patches give them."""
    phi = c.uniform(seed)
    P = phi[c.bFaceCells]
    fixed = c.bType == 1
    bPhi = np.where(fixed, c.bValue, P)
    bSnGrad = np.where(fixed, (c.bValue - P) * c.bDeltaCoeffs, 0.0)
    return phi, bPhi, bSnGrad


def c3_poisson_lognormal(n=256):
    """fvm::laplacian(gamma, p), gamma log-normal (sigma 1, correlation
    length 8 cells, seed 11), RHS seed 1: the benchmarked 256^3 system."""
    c = BoxCase(n)
    a = O.Addr(c.N, c.lower, c.upper)
    gamma = c.lognormal(11, 1.0, 8.0)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, c.weights, c.magSf, c.deltaCoeffs,
                                               c.bFaceCells, c.bMagSf, c.bDeltaCoeffs,
                                               c.bType, c.bValue)
    b = c.rhs(1)
    diag, b = O.add_boundary(c.bFaceCells, ic, bc, diag, b)
    return c, a, diag, upper, b


C3_CONTROLS = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-8, "relTol": 0,
               "maxIter": 20000}
C3_TIGHT = 1e-12
