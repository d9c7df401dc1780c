"""Synthetic hex-box cases (blockMesh ordering) for the measured
configurations, built by libfoamhip_case.so.

A BoxCase holds the LDU addressing, the uniform-box geometry and the
boundary description (patch 0 = x-min wall, fixedValue 0; the other five
patches zeroGradient).  Fields are seeded with the counter-based SplitMix64
generator of fh_case.cpp.
"""
import ctypes

import numpy as np

from . import _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class BoxCase:
    def __init__(self, nx, ny=None, nz=None, L=(1.0, 1.0, 1.0), fixed_patches=(0,)):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        lib = _lib.case()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny * nz
        self.F = int(lib.fhc_box_nfaces(nx, ny, nz))
        self.nB = int(lib.fhc_box_nbfaces(nx, ny, nz))
        self.lower = np.empty(self.F, np.int32)
        self.upper = np.empty(self.F, np.int32)
        self.faceDir = np.empty(self.F, np.int8)
        lib.fhc_box_faces(nx, ny, nz, _p(self.lower), _p(self.upper), _p(self.faceDir))
        self.bFaceCells = np.empty(self.nB, np.int32)
        self.bPatch = np.empty(self.nB, np.int8)
        lib.fhc_box_bfaces(nx, ny, nz, _p(self.bFaceCells), _p(self.bPatch))
        F, N, B = self.F, self.N, self.nB
        self.weights = np.empty(F)
        self.magSf = np.empty(F)
        self.Sf = np.empty(3 * F)
        self.deltaCoeffs = np.empty(F)
        self.V = np.empty(N)
        self.C = np.empty(3 * N)
        self.Cf = np.empty(3 * F)
        self.bMagSf = np.empty(B)
        self.bSf = np.empty(3 * B)
        self.bDeltaCoeffs = np.empty(B)
        self.bCf = np.empty(3 * B)
        lib.fhc_box_geometry(nx, ny, nz, L[0], L[1], L[2], _p(self.weights), _p(self.magSf),
                             _p(self.Sf), _p(self.deltaCoeffs), _p(self.V), _p(self.C),
                             _p(self.Cf), _p(self.bMagSf), _p(self.bSf),
                             _p(self.bDeltaCoeffs), _p(self.bCf))
        # nonOrthDeltaCoeffs == deltaCoeffs on an orthogonal uniform box
        self.nonOrthDeltaCoeffs = self.deltaCoeffs
        self.bType = np.isin(self.bPatch, np.asarray(fixed_patches, np.int8)).astype(np.int32)
        self.bValue = np.zeros(B)

    # seeded fields
    def rhs(self, seed):
        out = np.empty(self.N)
        _lib.case().fhc_rhs(seed, self.N, _p(out))
        return out

    def uniform(self, seed, lo=-1.0, hi=1.0, n=None):
        n = self.N if n is None else n
        out = np.empty(n)
        _lib.case().fhc_uniform(seed, n, lo, hi, _p(out))
        return out

    def lognormal(self, seed, sigma, corrLen=8.0):
        out = np.empty(self.N)
        _lib.case().fhc_lognormal(self.nx, self.ny, self.nz, seed, corrLen, sigma, _p(out))
        return out

    def fourier_flux(self, seed):
        phi = np.empty(self.F)
        bPhi = np.empty(self.nB)
        _lib.case().fhc_fourier_flux(self.nx, self.ny, self.nz, seed, _p(self.C), _p(self.Cf),
                                     _p(self.Sf), _p(self.bCf), _p(self.bSf), _p(phi),
                                     _p(bPhi))
        return phi, bPhi

    def laplacian_system(self, gamma, rhs):
        """Host assembly of fvm::laplacian(gamma, p) = rhs with the boundary
        coefficients added (as fvMatrix::solveSegregated does for the call):
        returns (diag, upper, source).  Workload generation for request
        streams; the device assembly is fh_assemble_laplacian."""
        gP, gN = gamma[self.lower], gamma[self.upper]
        upper = self.deltaCoeffs * ((self.weights * (gP - gN) + gN) * self.magSf)
        diag = -(np.bincount(self.lower, upper, self.N) + np.bincount(self.upper, upper, self.N))
        pGamma = gamma[self.bFaceCells] * self.bMagSf
        fixed = self.bType == 1
        ic = np.where(fixed, -pGamma * self.bDeltaCoeffs, 0.0)
        bc = np.where(fixed, -pGamma * (self.bDeltaCoeffs * self.bValue), 0.0)
        diag = diag + np.bincount(self.bFaceCells, ic, self.N)
        source = rhs + np.bincount(self.bFaceCells, bc, self.N)
        return diag, upper, source


class GradedBoxCase(BoxCase):
    """A box of nx x ny x nz hexes graded along z (expansion ratio zRatio,
    last cell over first) with its interior points jittered by `jitter` of
    the local cell size (seeded): the addressing and patches of BoxCase, the
    points and the quadrilateral faces (internal faces, then boundary faces
    in bFaceCells order, oriented owner -> neighbour / outwards).  The
    geometry (face centres and areas, cell centres and volumes, weights and
    delta coefficients) is computed from them, not assumed."""

    def __init__(self, nx, ny=None, nz=None, L=(1.0, 1.0, 1.0), zRatio=100.0, jitter=0.15,
                 seed=1, fixed_patches=(0,)):
        super().__init__(nx, ny, nz, L, fixed_patches)
        lib = _lib.case()
        nx, ny, nz = self.nx, self.ny, self.nz
        self.nPoints = (nx + 1) * (ny + 1) * (nz + 1)
        self.points = np.empty(3 * self.nPoints)
        lib.fhc_graded_box_points(nx, ny, nz, L[0], L[1], L[2], zRatio, jitter, seed,
                                  _p(self.points))
        nAll = self.F + self.nB
        self.facePoints = np.empty(4 * nAll, np.int32)
        lib.fhc_box_face_points(nx, ny, nz, _p(self.facePoints))
        self.faceStart = (4 * np.arange(nAll + 1)).astype(np.int32)
        # the uniform-box geometry of BoxCase does not describe this mesh
        for k in ("weights", "magSf", "Sf", "deltaCoeffs", "V", "C", "Cf", "bMagSf", "bSf",
                  "bDeltaCoeffs", "bCf", "nonOrthDeltaCoeffs"):
            setattr(self, k, None)
