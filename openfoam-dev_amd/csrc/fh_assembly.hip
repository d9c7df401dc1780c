// Finite-volume face loops that feed the solver, rewritten as gathers:
//   fvm::laplacian   gaussLaplacianScheme::fvmLaplacianUncorrected
//                    (gaussLaplacianScheme.C:45-89, gaussLaplacianSchemes.C:43-56)
//   fvm::div upwind  gaussConvectionScheme::fvmDiv (gaussConvectionScheme.C:119-170)
//   fvc::grad        gaussGrad::calcGrad (gaussGrad.C:30-75) + gradScheme.C:123-148
// Face values are computed once per face; each cell then sums its faces in
// the reference face order (neighbour faces first, then owner faces, then
// boundary faces in patch order), so the scattered reference loops are
// reproduced bitwise without atomics.
#include "fh_assembly.h"

namespace fh
{

#define FH_BLOCK 256

__device__ __forceinline__ int32_t sell_base3(const int32_t* sliceStart, int32_t r)
{
    return sliceStart[r >> 6] + (r & 63);
}

// upper[f] = deltaCoeffs*(gamma_f*magSf), gamma_f = w*(gP - gN) + gN
// (surfaceInterpolationScheme.C:305-314 with linear weights)
__global__ void k_lap_faces(int32_t F, const int32_t* __restrict__ lower,
                            const int32_t* __restrict__ upper, const double* __restrict__ gamma,
                            const double* __restrict__ w, const double* __restrict__ magSf,
                            const double* __restrict__ dc, double* __restrict__ upperOut)
{
    const int32_t f = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (f >= F) return;
    const double gP = gamma[lower[f]], gN = gamma[upper[f]];
    const double gf = w[f]*(gP - gN) + gN;
    upperOut[f] = dc[f]*(gf*magSf[f]);
}

// gaussConvectionScheme.C:140-142 with upwind weights pos0(phi)
// (upwind.H:116-128), minus the uniform-D laplacian coefficient
__global__ void k_convdiff_faces(int32_t F, const double* __restrict__ phi, double D,
                                 const double* __restrict__ magSf, const double* __restrict__ dc,
                                 double* __restrict__ upperOut, double* __restrict__ lowerOut,
                                 double* __restrict__ lapU, double* __restrict__ convL,
                                 double* __restrict__ convU)
{
    const int32_t f = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (f >= F) return;
    const double w = phi[f] >= 0.0 ? 1.0 : 0.0;
    const double lo = -w*phi[f];
    const double up = lo + phi[f];
    const double lu = dc[f]*(D*magSf[f]);
    convL[f] = lo;
    convU[f] = up;
    lapU[f] = lu;
    lowerOut[f] = lo - lu;
    upperOut[f] = up - lu;
}

// negSumDiag (lduMatrixOperations.C:55-74) as a per-row gather: the row is
// the neighbour of its lower entries (Diag[u] -= Upper[f]) and the owner of
// its upper entries (Diag[l] -= Lower[f]); then addBoundaryDiag
// (fvMatrix.C:112-127) over the row's boundary faces in patch order.
__device__ __forceinline__ double neg_sum_row(const RowsView& R, int32_t r,
                                              const double* __restrict__ upperF,
                                              const double* __restrict__ lowerF)
{
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = sell_base3(R.sliceStart, r);
    double acc = 0.0;
    for (int32_t k = 0; k < nEnt; k++)
    {
        const int32_t fc = R.face[base + (k << 6)];
        const int32_t f = fc >> 1;
        acc -= (fc & 1) ? lowerF[f] : upperF[f];
    }
    return acc;
}

__global__ void k_lap_cells(RowsView R, const int32_t* __restrict__ perm,
                            const double* __restrict__ upperF,
                            BoundaryView B, const double* __restrict__ ic,
                            int32_t addBoundaryDiag, double* __restrict__ diagLib)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    double d = neg_sum_row(R, r, upperF, upperF);
    if (addBoundaryDiag && B.nB > 0)
    {
        const int32_t c = perm[r];
        for (int32_t q = B.cellStart[c]; q < B.cellStart[c + 1]; q++) d += ic[B.cellFaces[q]];
    }
    diagLib[r] = d;
}

__global__ void k_convdiff_cells(RowsView R, const int32_t* __restrict__ perm,
                                 const double* __restrict__ convU, const double* __restrict__ convL,
                                 const double* __restrict__ lapU,
                                 BoundaryView B, const double* __restrict__ ic,
                                 int32_t addBoundaryDiag, double* __restrict__ diagLib)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    const double dConv = neg_sum_row(R, r, convU, convL);
    const double dLap = neg_sum_row(R, r, lapU, lapU);
    double d = dConv - dLap;
    if (addBoundaryDiag && B.nB > 0)
    {
        const int32_t c = perm[r];
        for (int32_t q = B.cellStart[c]; q < B.cellStart[c + 1]; q++) d += ic[B.cellFaces[q]];
    }
    diagLib[r] = d;
}

// boundary coefficients of the laplacian (gaussLaplacianScheme.C:69-86):
// fixedValue internal = gammaMagSf*(-deltaCoeffs), boundary =
// -gammaMagSf*(deltaCoeffs*value); zeroGradient 0.  The boundary gamma is
// the adjacent cell value.
__global__ void k_lap_bfaces(int32_t nB, const int32_t* __restrict__ bFaceCells,
                             const double* __restrict__ gamma, const double* __restrict__ bMagSf,
                             const double* __restrict__ bDc, const int32_t* __restrict__ bType,
                             const double* __restrict__ bValue, double* __restrict__ ic,
                             double* __restrict__ bc)
{
    const int32_t b = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (b >= nB) return;
    const double pGamma = gamma[bFaceCells[b]]*bMagSf[b];
    if (bType[b] == 1)
    {
        ic[b] = pGamma*(-bDc[b]);
        bc[b] = -pGamma*(bDc[b]*bValue[b]);
    }
    else
    {
        ic[b] = pGamma*0.0;
        bc[b] = -pGamma*0.0;
    }
}

// convection (fvmDiv :150-157) minus laplacian boundary coefficients
__global__ void k_convdiff_bfaces(int32_t nB, const double* __restrict__ bPhi, double D,
                                  const double* __restrict__ bMagSf,
                                  const double* __restrict__ bDc, const int32_t* __restrict__ bType,
                                  const double* __restrict__ bValue, double* __restrict__ ic,
                                  double* __restrict__ bc)
{
    const int32_t b = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (b >= nB) return;
    const double pGamma = D*bMagSf[b];
    double icC, bcC, icL, bcL;
    if (bType[b] == 1)
    {
        icC = bPhi[b]*0.0;
        bcC = -bPhi[b]*bValue[b];
        icL = pGamma*(-bDc[b]);
        bcL = -pGamma*(bDc[b]*bValue[b]);
    }
    else
    {
        icC = bPhi[b]*1.0;
        bcC = -bPhi[b]*0.0;
        icL = pGamma*0.0;
        bcL = -pGamma*0.0;
    }
    ic[b] = icC - icL;
    bc[b] = bcC - bcL;
}

// addBoundarySource (fvMatrix.C:146-175) on a natural-order source vector
__global__ void k_add_bsource(int32_t n, BoundaryView B, const double* __restrict__ bc,
                              double* __restrict__ source)
{
    const int32_t c = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (c >= n) return;
    double s = source[c];
    for (int32_t q = B.cellStart[c]; q < B.cellStart[c + 1]; q++) s += bc[B.cellFaces[q]];
    source[c] = s;
}

// gaussGrad.C:53-74: Sfsf = Sf*phi_f per face; grad[own] += Sfsf,
// grad[nei] -= Sfsf in face order; += boundary Sf*phi_b; /= V
__global__ void k_grad_faces(int32_t F, const int32_t* __restrict__ lower,
                             const int32_t* __restrict__ upper, const double* __restrict__ phi,
                             const double* __restrict__ w, const double* __restrict__ Sf,
                             double* __restrict__ Sfsf)
{
    const int32_t f = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (f >= F) return;
    const double pP = phi[lower[f]], pN = phi[upper[f]];
    const double sf = w[f]*(pP - pN) + pN;
    Sfsf[3*f] = Sf[3*f]*sf;
    Sfsf[3*f + 1] = Sf[3*f + 1]*sf;
    Sfsf[3*f + 2] = Sf[3*f + 2]*sf;
}

__global__ void k_grad_cells(RowsView R, const int32_t* __restrict__ perm,
                             const double* __restrict__ Sfsf, BoundaryView B,
                             const double* __restrict__ bSf, const double* __restrict__ bPhi,
                             const double* __restrict__ V, double* __restrict__ grad)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = sell_base3(R.sliceStart, r);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int32_t k = 0; k < nEnt; k++)
    {
        const int32_t fc = R.face[base + (k << 6)];
        const int32_t f = fc >> 1;
        if (fc & 1) { g0 += Sfsf[3*f]; g1 += Sfsf[3*f + 1]; g2 += Sfsf[3*f + 2]; }
        else        { g0 -= Sfsf[3*f]; g1 -= Sfsf[3*f + 1]; g2 -= Sfsf[3*f + 2]; }
    }
    const int32_t c = perm[r];
    if (B.nB > 0)
        for (int32_t q = B.cellStart[c]; q < B.cellStart[c + 1]; q++)
        {
            const int32_t b = B.cellFaces[q];
            g0 += bSf[3*b]*bPhi[b];
            g1 += bSf[3*b + 1]*bPhi[b];
            g2 += bSf[3*b + 2]*bPhi[b];
        }
    const double v = V[c];
    grad[3*c] = g0/v;
    grad[3*c + 1] = g1/v;
    grad[3*c + 2] = g2/v;
}

// gradScheme.C:123-148: patch gradient = cell gradient corrected to the
// patch-normal snGrad
__global__ void k_grad_bfaces(int32_t nB, const int32_t* __restrict__ bFaceCells,
                              const double* __restrict__ bSf, const double* __restrict__ bMagSf,
                              const double* __restrict__ bSnGrad, const double* __restrict__ grad,
                              double* __restrict__ bGrad)
{
    const int32_t b = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (b >= nB) return;
    const int32_t c = bFaceCells[b];
    const double n0 = bSf[3*b]/bMagSf[b], n1 = bSf[3*b + 1]/bMagSf[b], n2 = bSf[3*b + 2]/bMagSf[b];
    const double g0 = grad[3*c], g1 = grad[3*c + 1], g2 = grad[3*c + 2];
    const double ng = n0*g0 + n1*g1 + n2*g2;
    const double s = bSnGrad[b] - ng;
    bGrad[3*b] = g0 + n0*s;
    bGrad[3*b + 1] = g1 + n1*s;
    bGrad[3*b + 2] = g2 + n2*s;
}

static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + FH_BLOCK - 1)/FH_BLOCK)); }

void launch_laplacian(const RowsView& R, const int32_t* perm, const int32_t* lower,
                      const int32_t* upper, int32_t F, const BoundaryView& B,
                      const LaplacianInputs& in, double* upperF, double* diagLib,
                      double* ic, double* bc, double* source, int32_t addBoundaryDiag,
                      hipStream_t s)
{
    if (F > 0)
        hipLaunchKernelGGL(k_lap_faces, g1(F), dim3(FH_BLOCK), 0, s, F, lower, upper,
                           in.gamma, in.weights, in.magSf, in.deltaCoeffs, upperF);
    if (B.nB > 0)
        hipLaunchKernelGGL(k_lap_bfaces, g1(B.nB), dim3(FH_BLOCK), 0, s, B.nB, B.faceCells,
                           in.gamma, in.bMagSf, in.bDeltaCoeffs, in.bType, in.bValue, ic, bc);
    hipLaunchKernelGGL(k_lap_cells, g1(R.n), dim3(FH_BLOCK), 0, s, R, perm, upperF, B, ic,
                       addBoundaryDiag, diagLib);
    if (source && B.nB > 0)
        hipLaunchKernelGGL(k_add_bsource, g1(R.n), dim3(FH_BLOCK), 0, s, R.n, B, bc, source);
}

void launch_convection_diffusion(const RowsView& R, const int32_t* perm, int32_t F,
                                 const BoundaryView& B, const ConvDiffInputs& in,
                                 double* upperF, double* lowerF, double* work3F,
                                 double* diagLib, double* ic, double* bc, double* source,
                                 int32_t addBoundaryDiag, hipStream_t s)
{
    double* lapU = work3F;
    double* convL = work3F + F;
    double* convU = work3F + 2*(int64_t)F;
    if (F > 0)
        hipLaunchKernelGGL(k_convdiff_faces, g1(F), dim3(FH_BLOCK), 0, s, F, in.phi, in.D,
                           in.magSf, in.deltaCoeffs, upperF, lowerF, lapU, convL, convU);
    if (B.nB > 0)
        hipLaunchKernelGGL(k_convdiff_bfaces, g1(B.nB), dim3(FH_BLOCK), 0, s, B.nB, in.bPhi,
                           in.D, in.bMagSf, in.bDeltaCoeffs, in.bType, in.bValue, ic, bc);
    hipLaunchKernelGGL(k_convdiff_cells, g1(R.n), dim3(FH_BLOCK), 0, s, R, perm, convU, convL,
                       lapU, B, ic, addBoundaryDiag, diagLib);
    if (source && B.nB > 0)
        hipLaunchKernelGGL(k_add_bsource, g1(R.n), dim3(FH_BLOCK), 0, s, R.n, B, bc, source);
}

void launch_grad(const RowsView& R, const int32_t* perm, const int32_t* lower,
                 const int32_t* upper, int32_t F, const BoundaryView& B, const GradInputs& in,
                 double* work3F, double* grad, double* bGrad, hipStream_t s)
{
    if (F > 0)
        hipLaunchKernelGGL(k_grad_faces, g1(F), dim3(FH_BLOCK), 0, s, F, lower, upper, in.phi,
                           in.weights, in.Sf, work3F);
    hipLaunchKernelGGL(k_grad_cells, g1(R.n), dim3(FH_BLOCK), 0, s, R, perm, work3F, B, in.bSf,
                       in.bPhi, in.V, grad);
    if (bGrad && B.nB > 0)
        hipLaunchKernelGGL(k_grad_bfaces, g1(B.nB), dim3(FH_BLOCK), 0, s, B.nB, B.faceCells,
                           in.bSf, in.bMagSf, in.bSnGrad, grad, bGrad);
}

} // namespace fh
