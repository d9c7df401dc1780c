// lduMatrix H, H1, faceH and sumMagOffDiag: the off-diagonal face loops of
// the pressure-velocity coupling (the H operator of the momentum matrix,
// its coefficient sum H1, the face flux of H, and the off-diagonal
// magnitude sum that fvMatrix::relax starts from).
//
// The reference walks the faces once and scatters into both cells.  Here
// each library row is one thread over its SELL-64 entries: the lower-
// neighbour entries by ascending face, then the upper-neighbour entries by
// ascending face, which is the order in which the face loop reaches the
// cell, so every sum is the reference's to the bit.  The row's coefficient
// of a neighbour is the one the face loop uses for that cell (the lower
// coefficient for the neighbour cell of a face, the upper one for its
// owner), i.e. the SELL values of the product A x.
#include "fh_hops.h"

namespace fh
{

#define HO_B 256

// entry k of row r at sliceStart[r/64] + 64*k + r%64
__device__ __forceinline__ int32_t ho_base(const int32_t* sliceStart, int32_t r)
{
    return sliceStart[r >> 6] + (r & 63);
}

static dim3 grid_of(int64_t n) { return dim3((unsigned)((n + HO_B - 1)/HO_B)); }

// lduMatrixTemplates.C:33-64 for NC components (NC = 0: nC at run time,
// one component per launch, component c0)
template<int NC>
__global__ void k_h(RowsView R, const double* __restrict__ val, int32_t nC, int32_t c0,
                    const double* __restrict__ x, double* __restrict__ y)
{
    const int32_t r = blockIdx.x*HO_B + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = ho_base(R.sliceStart, r);
    if constexpr (NC > 0)
    {
        double acc[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] = 0.0;
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            const double a = val[e];
            const int64_t j = (int64_t)R.col[e]*NC;
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] -= a*x[j + c];
        }
#pragma unroll
        for (int c = 0; c < NC; c++) y[(int64_t)r*NC + c] = acc[c];
    }
    else
    {
        double acc = 0.0;
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            acc -= val[e]*x[(int64_t)R.col[e]*nC + c0];
        }
        y[(int64_t)r*nC + c0] = acc;
    }
}

// lduMatrixATmul.C:298-324
__global__ void k_h1(RowsView R, const double* __restrict__ val, double* __restrict__ y)
{
    const int32_t r = blockIdx.x*HO_B + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = ho_base(R.sliceStart, r);
    double acc = 0.0;
    for (int32_t k = 0; k < nEnt; k++) acc -= val[base + (k << 6)];
    y[r] = acc;
}

// lduMatrixOperations.C:77-97 (added to y)
__global__ void k_sum_mag_off_diag(RowsView R, const double* __restrict__ val,
                                   double* __restrict__ y)
{
    const int32_t r = blockIdx.x*HO_B + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = ho_base(R.sliceStart, r);
    double acc = y[r];
    for (int32_t k = 0; k < nEnt; k++) acc += fabs(val[base + (k << 6)]);
    y[r] = acc;
}

// lduMatrixTemplates.C:77-111: one thread per face, natural numbering
__global__ void k_face_h(int32_t nFaces, const int32_t* __restrict__ l,
                         const int32_t* __restrict__ u, const double* __restrict__ upper,
                         const double* __restrict__ lower, int32_t nC,
                         const double* __restrict__ x, double* __restrict__ y)
{
    const int32_t f = blockIdx.x*HO_B + threadIdx.x;
    if (f >= nFaces) return;
    const double U = upper[f], L = lower[f];
    const int64_t jl = (int64_t)l[f]*nC, ju = (int64_t)u[f]*nC;
    for (int32_t c = 0; c < nC; c++) y[(int64_t)f*nC + c] = U*x[ju + c] - L*x[jl + c];
}

// fvMatrix.C:594-740, fvMatrix<Type>::relax(alpha), every patch
// non-coupled: per row the off-diagonal magnitude sum (sumMagOffDiag from
// zero, row order = face order), the cell's boundary faces in patch order
// (the largest internal-coefficient magnitude added, then after the
// dominance and relaxation the smallest coefficient taken off), and the
// source update S += (D - D0) psi.  max and min select as the reference's
// (s1 > s2 ? s1 : s2, s1 < s2 ? s1 : s2).
__global__ void k_relax(RowsView R, const int32_t* __restrict__ perm,
                        const double* __restrict__ val, BoundaryView B,
                        const double* __restrict__ ic, int32_t nC, double alpha,
                        const double* __restrict__ psi, double* __restrict__ diag,
                        double* __restrict__ source, double* __restrict__ diagOut)
{
    const int32_t r = blockIdx.x*HO_B + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = ho_base(R.sliceStart, r);
    double sumOff = 0.0;
    for (int32_t k = 0; k < nEnt; k++) sumOff += fabs(val[base + (k << 6)]);
    const int32_t c = perm[r];
    const double D0 = diag[r];
    double D = D0;
    const int32_t q0 = B.nB > 0 ? B.cellStart[c] : 0, q1 = B.nB > 0 ? B.cellStart[c + 1] : 0;
    for (int32_t q = q0; q < q1; q++)
    {
        const double* v = ic + (int64_t)B.cellFaces[q]*nC;
        double mx = fabs(v[0]);
        for (int32_t k = 1; k < nC; k++)
        {
            const double a = fabs(v[k]);
            mx = mx > a ? mx : a;
        }
        D += mx;
    }
    const double aD = fabs(D);
    D = aD > sumOff ? aD : sumOff;
    D /= alpha;
    for (int32_t q = q0; q < q1; q++)
    {
        const double* v = ic + (int64_t)B.cellFaces[q]*nC;
        double mn = v[0];
        for (int32_t k = 1; k < nC; k++) mn = mn < v[k] ? mn : v[k];
        D -= mn;
    }
    diag[r] = D;
    if (diagOut) diagOut[c] = D;
    const double dD = D - D0;
    for (int32_t k = 0; k < nC; k++)
    {
        const int64_t i = (int64_t)c*nC + k;
        source[i] += dD*psi[i];
    }
}

void launch_relax(const RowsView& R, const int32_t* perm, const double* val,
                  const BoundaryView& B, const double* ic, int32_t nC, double alpha,
                  const double* psi, double* diag, double* source, double* diagOut,
                  hipStream_t s)
{
    if (R.n == 0) return;
    hipLaunchKernelGGL(k_relax, dim3((unsigned)((R.n + HO_B - 1)/HO_B)), dim3(HO_B), 0, s, R,
                       perm, val, B, ic, nC, alpha, psi, diag, source, diagOut);
}

// fvMatrix.C:953-1028, fvMatrix<Type>::flux() with non-coupled patches:
// internal faces faceH (+ the face flux correction), boundary faces
// cmptMultiply(internalCoeffs, patchInternalField) - boundaryCoeffs
// (+ its boundary part); one thread per face
__global__ void k_flux_faces(int32_t nFaces, const int32_t* __restrict__ l,
                             const int32_t* __restrict__ u, const double* __restrict__ upper,
                             const double* __restrict__ lower, int32_t nC,
                             const double* __restrict__ x, const double* __restrict__ corr,
                             double* __restrict__ y)
{
    const int32_t f = blockIdx.x*HO_B + threadIdx.x;
    if (f >= nFaces) return;
    const double U = upper[f], L = lower[f];
    const int64_t jl = (int64_t)l[f]*nC, ju = (int64_t)u[f]*nC;
    for (int32_t c = 0; c < nC; c++)
    {
        const int64_t i = (int64_t)f*nC + c;
        const double h = U*x[ju + c] - L*x[jl + c];
        y[i] = corr ? h + corr[i] : h;
    }
}

__global__ void k_flux_bfaces(int32_t nB, const int32_t* __restrict__ faceCells, int32_t nC,
                              const double* __restrict__ x, const double* __restrict__ ic,
                              const double* __restrict__ bc, const double* __restrict__ corr,
                              double* __restrict__ y)
{
    const int32_t b = blockIdx.x*HO_B + threadIdx.x;
    if (b >= nB) return;
    const int64_t j = (int64_t)faceCells[b]*nC;
    for (int32_t c = 0; c < nC; c++)
    {
        const int64_t i = (int64_t)b*nC + c;
        const double v = ic[i]*x[j + c] - bc[i];
        y[i] = corr ? v + corr[i] : v;
    }
}

void launch_flux(int32_t nFaces, const int32_t* l, const int32_t* u, const double* upper,
                 const double* lower, const BoundaryView& B, int32_t nC, const double* x,
                 const double* ic, const double* bc, const double* corr, const double* bCorr,
                 double* y, double* by, hipStream_t s)
{
    if (nFaces > 0)
        hipLaunchKernelGGL(k_flux_faces, grid_of(nFaces), dim3(HO_B), 0, s, nFaces, l, u, upper,
                           lower, nC, x, corr, y);
    if (B.nB > 0)
        hipLaunchKernelGGL(k_flux_bfaces, grid_of(B.nB), dim3(HO_B), 0, s, B.nB, B.faceCells,
                           nC, x, ic, bc, bCorr, by);
}

__global__ void k_permute_cmpt(int32_t n, int32_t nC, const int32_t* __restrict__ perm,
                               const double* __restrict__ in, double* __restrict__ out,
                               bool toLib)
{
    const int64_t i = blockIdx.x*(int64_t)HO_B + threadIdx.x;
    if (i >= (int64_t)n*nC) return;
    const int32_t r = (int32_t)(i/nC), c = (int32_t)(i % nC);
    const int64_t nat = (int64_t)perm[r]*nC + c;
    if (toLib) out[i] = in[nat];
    else out[nat] = in[i];
}

void launch_h(const RowsView& R, const double* val, int32_t nC, const double* x, double* y,
              hipStream_t s)
{
    if (R.n == 0) return;
    if (nC == 1)
        hipLaunchKernelGGL(k_h<1>, grid_of(R.n), dim3(HO_B), 0, s, R, val, 1, 0, x, y);
    else if (nC == 3)
        hipLaunchKernelGGL(k_h<3>, grid_of(R.n), dim3(HO_B), 0, s, R, val, 3, 0, x, y);
    else
        for (int32_t c = 0; c < nC; c++)
            hipLaunchKernelGGL(k_h<0>, grid_of(R.n), dim3(HO_B), 0, s, R, val, nC, c, x,
                               y);
}

void launch_h1(const RowsView& R, const double* val, double* y, hipStream_t s)
{
    if (R.n == 0) return;
    hipLaunchKernelGGL(k_h1, grid_of(R.n), dim3(HO_B), 0, s, R, val, y);
}

void launch_sum_mag_off_diag(const RowsView& R, const double* val, double* y, hipStream_t s)
{
    if (R.n == 0) return;
    hipLaunchKernelGGL(k_sum_mag_off_diag, grid_of(R.n), dim3(HO_B), 0, s, R, val, y);
}

void launch_face_h(int32_t nFaces, const int32_t* l, const int32_t* u, const double* upper,
                   const double* lower, int32_t nC, const double* x, double* y, hipStream_t s)
{
    if (nFaces == 0) return;
    hipLaunchKernelGGL(k_face_h, grid_of(nFaces), dim3(HO_B), 0, s, nFaces, l, u, upper,
                       lower, nC, x, y);
}

void launch_permute_cmpt_in(int32_t n, int32_t nC, const int32_t* perm, const double* xNat,
                            double* xLib, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_permute_cmpt, grid_of((int64_t)n*nC), dim3(HO_B), 0, s, n, nC,
                       perm, xNat, xLib, true);
}

void launch_permute_cmpt_out(int32_t n, int32_t nC, const int32_t* perm, const double* xLib,
                             double* xNat, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_permute_cmpt, grid_of((int64_t)n*nC), dim3(HO_B), 0, s, n, nC,
                       perm, xLib, xNat, false);
}

} // namespace fh
