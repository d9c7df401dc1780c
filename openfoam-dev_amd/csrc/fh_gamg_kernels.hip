// Device kernels of GAMG: restriction / prolongation gathers, Galerkin
// coarse coefficients, correction scaling and the dense coarsest solve.
#include "fh_gamg_kernels.h"

#include <cfloat>

namespace fh
{

#define GB 256

static inline dim3 gg(int64_t n) { return dim3((unsigned)((n + GB - 1)/GB)); }

// restrictField (GAMGAgglomerationTemplates.C:76-90): each coarse cell sums
// its fine cells in ascending fine cell order, starting from zero
template<class T>
__global__ void k_restrict(int32_t nc, const int32_t* __restrict__ start,
                           const int32_t* __restrict__ cells, const T* __restrict__ ff,
                           T* __restrict__ cf, const DevState* st)
{
    if (st && st->done) return;
    const int32_t c = blockIdx.x*GB + threadIdx.x;
    if (c >= nc) return;
    T s = T(0);
    for (int32_t q = start[c]; q < start[c + 1]; q++) s += ff[cells[q]];
    cf[c] = s;
}

// prolongField (GAMGAgglomerationTemplates.C:172-219): injection
template<class T>
__global__ void k_prolong(int32_t nf, const int32_t* __restrict__ map, const T* __restrict__ cf,
                          T* __restrict__ ff, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < nf) ff[i] = cf[map[i]];
}

// GAMGSolverAgglomerateMatrix.C:76-193: coarse diagonal = sum of the fine
// diagonals (fine cell order) + intra-cluster faces (fine face order),
// 2*upper (symmetric) or upper+lower
__global__ void k_galerkin_diag(int32_t nc, const int32_t* __restrict__ dStart,
                                const int32_t* __restrict__ dCells,
                                const int32_t* __restrict__ iStart,
                                const int32_t* __restrict__ iFaces,
                                const double* __restrict__ fineDiagLib,
                                const double* __restrict__ fineUpper,
                                const double* __restrict__ fineLower,
                                double* __restrict__ coarseDiagLib)
{
    const int32_t c = blockIdx.x*GB + threadIdx.x;
    if (c >= nc) return;
    double s = 0.0;
    for (int32_t q = dStart[c]; q < dStart[c + 1]; q++) s += fineDiagLib[dCells[q]];
    if (fineLower)
        for (int32_t q = iStart[c]; q < iStart[c + 1]; q++)
        {
            const int32_t f = iFaces[q];
            s += fineUpper[f] + fineLower[f];
        }
    else
        for (int32_t q = iStart[c]; q < iStart[c + 1]; q++) s += 2*fineUpper[iFaces[q]];
    coarseDiagLib[c] = s;
}

// coarse face coefficients: fine faces mapped to the coarse face, in fine
// face order, swapped when the fine face is flipped
__global__ void k_galerkin_faces(int32_t ncf, const int32_t* __restrict__ fStart,
                                 const int32_t* __restrict__ fFaces,
                                 const uint8_t* __restrict__ fFlip,
                                 const double* __restrict__ fineUpper,
                                 const double* __restrict__ fineLower,
                                 double* __restrict__ cu, double* __restrict__ cl)
{
    const int32_t f = blockIdx.x*GB + threadIdx.x;
    if (f >= ncf) return;
    double u = 0.0, l = 0.0;
    if (fineLower)
    {
        for (int32_t q = fStart[f]; q < fStart[f + 1]; q++)
        {
            const int32_t ff = fFaces[q];
            if (!fFlip[q]) { u += fineUpper[ff]; l += fineLower[ff]; }
            else { u += fineLower[ff]; l += fineUpper[ff]; }
        }
        cl[f] = l;
    }
    else
    {
        for (int32_t q = fStart[f]; q < fStart[f + 1]; q++) u += fineUpper[fFaces[q]];
    }
    cu[f] = u;
}

// scale (GAMGSolverScale.C:31-76): Acf = A field with the two sums
// sum source*field and sum Acf*field
template<class T>
__global__ void k_scale_spmv(RowsView R, const T* __restrict__ diag, const T* __restrict__ val,
                             const T* __restrict__ field, const T* __restrict__ source,
                             T* __restrict__ Acf, double* partials, const DevState* st)
{
    if (st && st->done) return;
    __shared__ double sm[2*GB/64];
    const int32_t r = blockIdx.x*GB + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (r < R.n)
    {
        const int32_t nEnt = R.rowInfo[r] >> 16;
        const int32_t base = R.sliceStart[r >> 6] + (r & 63);
        T acc = diag[r]*field[r];
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            acc += val[e]*field[R.col[e]];
        }
        Acf[r] = acc;
        a = (double)source[r]*(double)field[r];
        b = (double)acc*(double)field[r];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { sm[wid] = a; sm[GB/64 + wid] = b; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < GB/64; i++) { sa += sm[i]; sb += sm[GB/64 + i]; }
        partials[blockIdx.x] = sa;
        partials[gridDim.x + blockIdx.x] = sb;
    }
}

__global__ void k_scale_factor(int32_t nBlocks, const double* partials, DevState* st)
{
    if (st->done) return;
    __shared__ double sm[2*GB/64];
    double a = 0.0, b = 0.0;
    for (int32_t i = threadIdx.x; i < nBlocks; i += GB)
    {
        a += partials[i];
        b += partials[nBlocks + i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { sm[wid] = a; sm[GB/64 + wid] = b; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < GB/64; i++) { sa += sm[i]; sb += sm[GB/64 + i]; }
        // stabilise(den, vSmall) (Scalar.H:367-377)
        const double den = sb >= 0 ? sb + DBL_MIN : sb - DBL_MIN;
        st->gamgScale = sa/den;
    }
}

template<class T>
__global__ void k_scale_apply(int32_t n, T* __restrict__ field, const T* __restrict__ Acf,
                              const T* __restrict__ source, const T* __restrict__ diag,
                              const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i >= n) return;
    const T sf = (T)st->gamgScale;
    field[i] = sf*field[i] + (source[i] - sf*Acf[i])/diag[i];
}

// dense copy of the coarsest ldu matrix (LUscalarMatrix::convert), column-
// major n x n, library row order of the coarsest mesh
__global__ void k_dense_fill(int32_t n, RowsView R, const double* __restrict__ diag,
                             const double* __restrict__ val, double* __restrict__ A)
{
    const int32_t r = blockIdx.x*GB + threadIdx.x;
    if (r >= n) return;
    A[(int64_t)r*n + r] = diag[r];
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = R.sliceStart[r >> 6] + (r & 63);
    for (int32_t k = 0; k < nEnt; k++)
    {
        const int32_t e = base + (k << 6);
        A[(int64_t)R.col[e]*n + r] = val[e];     // column-major: A(r, col)
    }
}

// Newton-Schulz helpers: T = 2I - T; and ||I - A X||_max partial maxima
__global__ void k_two_minus(int32_t n, double* __restrict__ Tm)
{
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    if (i >= (int64_t)n*n) return;
    const int64_t r = i % n, c = i/n;
    Tm[i] = (r == c ? 2.0 : 0.0) - Tm[i];
}

__global__ void k_ident_err(int32_t n, const double* __restrict__ AX, double* partMax)
{
    __shared__ double sm[GB];
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    double e = 0.0;
    if (i < (int64_t)n*n)
    {
        const int64_t r = i % n, c = i/n;
        e = fabs((r == c ? 1.0 : 0.0) - AX[i]);
    }
    sm[threadIdx.x] = e;
    __syncthreads();
    for (int s = GB/2; s > 0; s >>= 1)
    {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partMax[blockIdx.x] = sm[0];
}

__global__ void k_scaled_transpose(int32_t n, const double* __restrict__ A, double scale,
                                   double* __restrict__ X)
{
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    if (i >= (int64_t)n*n) return;
    const int64_t r = i % n, c = i/n;
    X[i] = scale*A[r*n + c];
}

// coarsest correction x = X b: one thread per row, columns in ascending
// order (fixed summation order, coalesced over the rows of a column)
template<class T>
__global__ void k_gemv(int32_t n, const double* __restrict__ X, const T* __restrict__ b,
                       T* __restrict__ x, const DevState* st)
{
    if (st && st->done) return;
    const int32_t r = blockIdx.x*GB + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int32_t c = 0; c < n; c++) s += X[(int64_t)c*n + r]*(double)b[c];
    x[r] = (T)s;
}

template<class T>
__global__ void k_axpy_gamg(int32_t n, T* __restrict__ y, const T* __restrict__ x,
                            const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) y[i] += x[i];
}

template<class T>
__global__ void k_sub_gamg(int32_t n, T* __restrict__ y, const T* __restrict__ x,
                           const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) y[i] -= x[i];
}

template<class T>
__global__ void k_diff_gamg(int32_t n, const T* __restrict__ a, const T* __restrict__ b,
                            T* __restrict__ out, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

template<class T>
__global__ void k_copy_gamg(int32_t n, const T* __restrict__ a, T* __restrict__ b,
                            const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) b[i] = a[i];
}

template<class T>
__global__ void k_zero_gamg(int32_t n, T* __restrict__ a, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) a[i] = T(0);
}

// residual scale of the mixed GAMG cycle: mean |r| from the last FP64
// residual sum (finalResidual*normFactor), 1 when that is not usable
__device__ __forceinline__ double mixed_scale(int32_t n, const DevState* st)
{
    const double sc = st->finalResidual*st->normFactor/n;
    return (sc > 0.0 && isfinite(sc) && isfinite(1.0/sc)) ? sc : 1.0;
}

__global__ void k_to_f32_gamg(int32_t n, const double* __restrict__ a, float* __restrict__ b,
                              const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) b[i] = (float)(a[i]*(1.0/mixed_scale(n, st)));
}

__global__ void k_add_f32_gamg(int32_t n, double* __restrict__ y, const float* __restrict__ x,
                               const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*GB + threadIdx.x;
    if (i < n) y[i] += mixed_scale(n, st)*(double)x[i];
}

// ---------------------------------------------------------------------------
// coarsest level, iterative: one 64-lane workgroup runs the whole Krylov
// solve of GAMGSolver::solveCoarsestLevel (GAMGSolverSolve.C:520-552 with
// the solvers of GAMGSolver.C:270-322): PCG+DIC for a symmetric matrix,
// PBiCGStab+DILU otherwise, from a zero initial guess, with the parent's
// tolerance and relTol.  Natural numbering; SpMV rows and point-wise
// updates spread over the lanes (per-row sums in the reference order), the
// triangular sweeps walk the faces on lane 0 exactly as the reference
// loops, reductions are fixed-order wave sums.
// ---------------------------------------------------------------------------
#define CW 64

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// setup: natural diag, reciprocal DIC / DILU factor (DICPreconditioner.C:57-84,
// DILUPreconditioner.C:57-85) and sumA (lduMatrixATmul.C:154-200)
__global__ void k_coarsest_setup(CoarsestSys S)
{
    if (threadIdx.x != 0) return;
    const int32_t n = S.n;
    for (int32_t i = 0; i < n; i++) S.diagN[S.perm[i]] = S.diagLib[i];
    for (int32_t c = 0; c < n; c++) { S.rD[c] = S.diagN[c]; S.sumA[c] = S.diagN[c]; }
    for (int32_t f = 0; f < S.nF; f++)
    {
        const int32_t l = S.lower[f], u = S.upper[f];
        S.rD[u] -= S.upperF[f]*S.lowerF[f]/S.rD[l];
        S.sumA[u] += S.lowerF[f];
        S.sumA[l] += S.upperF[f];
    }
    for (int32_t c = 0; c < n; c++) S.rD[c] = 1.0/S.rD[c];
}

__device__ __forceinline__ void cs_amul(const CoarsestSys& S, const double* diagN,
                                        const double* x, double* y)
{
    for (int32_t i = threadIdx.x; i < S.n; i += CW)
    {
        double acc = diagN[i]*x[i];
        for (int32_t q = S.rowStart[i]; q < S.rowStart[i + 1]; q++)
        {
            const int32_t e = S.ent[q];
            const double a = (e & 1) ? S.upperF[e >> 1] : S.lowerF[e >> 1];
            acc += a*x[S.colN[q]];
        }
        y[i] = acc;
    }
    __syncthreads();
}

// DIC (symmetric) / DILU precondition; the DILU forward sweep in face order
// equals the losort walk row by row
__device__ __forceinline__ void cs_precond(const CoarsestSys& S, const double* rD,
                                          const double* r, double* w)
{
    for (int32_t i = threadIdx.x; i < S.n; i += CW) w[i] = rD[i]*r[i];
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int32_t f = 0; f < S.nF; f++)
        {
            const int32_t l = S.lower[f], u = S.upper[f];
            w[u] -= rD[u]*S.lowerF[f]*w[l];
        }
        for (int32_t f = S.nF - 1; f >= 0; f--)
        {
            const int32_t l = S.lower[f], u = S.upper[f];
            w[l] -= rD[l]*S.upperF[f]*w[u];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double cs_dot(int32_t n, const double* a, const double* b)
{
    double s = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += CW) s += a[i]*b[i];
    return wave_sum(s);
}

__device__ __forceinline__ double cs_mag(int32_t n, const double* a)
{
    double s = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += CW) s += fabs(a[i]);
    return wave_sum(s);
}

template<class T>
__global__ void __launch_bounds__(CW) k_coarsest_krylov(CoarsestSys S, const T* srcLib,
                                                        T* corrLib, const DevState* st,
                                                        int useLds)
{
    if (st && st->done) return;
    extern __shared__ double lds[];
    const int32_t n = S.n;
    double* W = useLds ? lds : S.work;
    double* diagN = W;            // copies of the setup arrays
    double* rD = W + n;
    double* sumA = W + 2*n;
    double* psi = W + 3*n;
    double* src = W + 4*n;
    double* rA = W + 5*n;
    double* wA = W + 6*n;         // PCG wA / PBiCGStab yA
    double* pA = W + 7*n;
    double* AyA = W + 8*n;
    double* sA = W + 9*n;
    double* zA = W + 10*n;
    double* tA = W + 11*n;
    double* rA0 = W + 12*n;
    for (int32_t i = threadIdx.x; i < n; i += CW)
    {
        diagN[i] = S.diagN[i];
        rD[i] = S.rD[i];
        sumA[i] = S.sumA[i];
        psi[i] = 0.0;
        src[S.perm[i]] = (double)srcLib[i];
    }
    __syncthreads();

    // lduMatrixSolver.C:174-197 with psi = 0
    cs_amul(S, diagN, psi, wA);
    for (int32_t i = threadIdx.x; i < n; i += CW) rA[i] = src[i] - wA[i];
    double xs = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += CW) xs += psi[i];
    const double xRef = wave_sum(xs)/n;
    double nfs = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += CW)
    {
        const double t = sumA[i]*xRef;
        nfs += fabs(wA[i] - t) + fabs(src[i] - t);
    }
    const double normFactor = wave_sum(nfs) + 1e-20;
    __syncthreads();
    const double initial = cs_mag(n, rA)/normFactor;
    double fin = initial;
    auto converged = [&](double r)
    {
        return r < S.tol || (S.relTol > 1e-20 && r < S.relTol*initial);
    };
    int32_t nIter = 0;
    if (S.minIter > 0 || !converged(fin))
    {
        if (S.sym)
        {
            // PCG.C:129-189
            double wArA = 1e20, wArAold;
            do
            {
                wArAold = wArA;
                cs_precond(S, rD, rA, wA);
                wArA = cs_dot(n, wA, rA);
                if (nIter == 0)
                    for (int32_t i = threadIdx.x; i < n; i += CW) pA[i] = wA[i];
                else
                {
                    const double beta = wArA/wArAold;
                    for (int32_t i = threadIdx.x; i < n; i += CW) pA[i] = wA[i] + beta*pA[i];
                }
                __syncthreads();
                cs_amul(S, diagN, pA, wA);
                const double wApA = cs_dot(n, wA, pA);
                if (fabs(wApA)/normFactor < DBL_MIN) break;
                const double alpha = wArA/wApA;
                for (int32_t i = threadIdx.x; i < n; i += CW)
                {
                    psi[i] += alpha*pA[i];
                    rA[i] -= alpha*wA[i];
                }
                __syncthreads();
                fin = cs_mag(n, rA)/normFactor;
            } while ((++nIter < S.maxIter && !converged(fin)) || nIter < S.minIter);
        }
        else
        {
            // PBiCGStab.C:118-250
            double* yA = wA;
            for (int32_t i = threadIdx.x; i < n; i += CW) rA0[i] = rA[i];
            __syncthreads();
            double rA0rA = 0.0, alpha = 0.0, omega = 0.0;
            do
            {
                const double rA0rAold = rA0rA;
                rA0rA = cs_dot(n, rA0, rA);
                if (fabs(rA0rA) < DBL_MIN) break;
                if (nIter == 0)
                    for (int32_t i = threadIdx.x; i < n; i += CW) pA[i] = rA[i];
                else
                {
                    if (fabs(omega) < DBL_MIN) break;
                    const double beta = (rA0rA/rA0rAold)*(alpha/omega);
                    for (int32_t i = threadIdx.x; i < n; i += CW)
                        pA[i] = rA[i] + beta*(pA[i] - omega*AyA[i]);
                }
                __syncthreads();
                cs_precond(S, rD, pA, yA);
                cs_amul(S, diagN, yA, AyA);
                const double rA0AyA = cs_dot(n, rA0, AyA);
                alpha = rA0rA/rA0AyA;
                for (int32_t i = threadIdx.x; i < n; i += CW) sA[i] = rA[i] - alpha*AyA[i];
                __syncthreads();
                fin = cs_mag(n, sA)/normFactor;
                if (++nIter >= S.minIter && converged(fin))
                {
                    for (int32_t i = threadIdx.x; i < n; i += CW) psi[i] += alpha*yA[i];
                    __syncthreads();
                    break;
                }
                cs_precond(S, rD, sA, zA);
                cs_amul(S, diagN, zA, tA);
                const double tAtA = cs_dot(n, tA, tA);
                omega = cs_dot(n, tA, sA)/tAtA;
                for (int32_t i = threadIdx.x; i < n; i += CW)
                {
                    psi[i] += alpha*yA[i] + omega*zA[i];
                    rA[i] = sA[i] - omega*tA[i];
                }
                __syncthreads();
                fin = cs_mag(n, rA)/normFactor;
            } while ((nIter < S.maxIter && !converged(fin)) || nIter < S.minIter);
        }
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < n; i += CW) corrLib[i] = (T)psi[S.perm[i]];
}

// LUscalarMatrix::solve -> LUBacksubstitute (scalarMatricesTemplates.C:117-162)
// on lane 0 in natural order; LU and pivots come from the host Crout
// decomposition (scalarMatrices.C:77-180)
template<class T>
__global__ void k_lu_backsub(int32_t n, const double* __restrict__ LU,
                             const int32_t* __restrict__ pivot, const int32_t* __restrict__ perm,
                             double* __restrict__ x, const T* __restrict__ srcLib,
                             T* __restrict__ corrLib, const DevState* st)
{
    if (st && st->done) return;
    for (int32_t i = threadIdx.x; i < n; i += CW) x[perm[i]] = (double)srcLib[i];
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int32_t ii = 0;
        for (int32_t i = 0; i < n; i++)
        {
            const int32_t ip = pivot[i];
            double sum = x[ip];
            x[ip] = x[i];
            const double* Li = LU + (int64_t)i*n;
            if (ii != 0)
                for (int32_t j = ii - 1; j < i; j++) sum -= Li[j]*x[j];
            else if (sum != 0.0)
                ii = i + 1;
            x[i] = sum;
        }
        for (int32_t i = n - 1; i >= 0; i--)
        {
            double sum = x[i];
            const double* Li = LU + (int64_t)i*n;
            for (int32_t j = i + 1; j < n; j++) sum -= Li[j]*x[j];
            x[i] = sum/Li[i];
        }
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < n; i += CW) corrLib[i] = (T)x[perm[i]];
}

__global__ void k_fill_zero_d(int64_t n, double* a)
{
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ---- launchers ------------------------------------------------------------------

template<class T>
void launch_restrict(int32_t nc, const int32_t* start, const int32_t* cells, const T* ff,
                     T* cf, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_restrict<T>, gg(nc), dim3(GB), 0, s, nc, start, cells, ff, cf, st);
}

template<class T>
void launch_prolong(int32_t nf, const int32_t* map, const T* cf, T* ff, const DevState* st,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_prolong<T>, gg(nf), dim3(GB), 0, s, nf, map, cf, ff, st);
}

void launch_galerkin(int32_t nc, int32_t ncf, const GalerkinMaps& G, const double* fineDiagLib,
                     const double* fineUpper, const double* fineLower, double* coarseDiagLib,
                     double* cu, double* cl, hipStream_t s)
{
    hipLaunchKernelGGL(k_galerkin_diag, gg(nc), dim3(GB), 0, s, nc, G.dStart, G.dCells,
                       G.iStart, G.iFaces, fineDiagLib, fineUpper, fineLower, coarseDiagLib);
    if (ncf > 0)
        hipLaunchKernelGGL(k_galerkin_faces, gg(ncf), dim3(GB), 0, s, ncf, G.fStart, G.fFaces,
                           G.fFlip, fineUpper, fineLower, cu, cl);
}

template<class T>
void launch_gamg_scale(const RowsView& R, const T* diag, const T* val, T* field, T* Acf,
                       const T* source, double* partials, DevState* st, hipStream_t s)
{
    const int32_t nb = (R.n + GB - 1)/GB;
    hipLaunchKernelGGL(k_scale_spmv<T>, dim3(nb), dim3(GB), 0, s, R, diag, val, field, source,
                       Acf, partials, st);
    hipLaunchKernelGGL(k_scale_factor, dim3(1), dim3(GB), 0, s, nb, partials, st);
    hipLaunchKernelGGL(k_scale_apply<T>, gg(R.n), dim3(GB), 0, s, R.n, field, Acf, source,
                       diag, st);
}

void launch_dense_fill(int32_t n, const RowsView& R, const double* diag, const double* val,
                       double* A, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_zero_d, gg((int64_t)n*n), dim3(GB), 0, s, (int64_t)n*n, A);
    hipLaunchKernelGGL(k_dense_fill, gg(n), dim3(GB), 0, s, n, R, diag, val, A);
}

void launch_two_minus(int32_t n, double* T_, hipStream_t s)
{
    hipLaunchKernelGGL(k_two_minus, gg((int64_t)n*n), dim3(GB), 0, s, n, T_);
}

int64_t launch_ident_err(int32_t n, const double* AX, double* partMax, hipStream_t s)
{
    const int64_t nb = ((int64_t)n*n + GB - 1)/GB;
    hipLaunchKernelGGL(k_ident_err, dim3((unsigned)nb), dim3(GB), 0, s, n, AX, partMax);
    return nb;
}

void launch_scaled_transpose(int32_t n, const double* A, double scale, double* X,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_scaled_transpose, gg((int64_t)n*n), dim3(GB), 0, s, n, A, scale, X);
}

template<class T>
void launch_gemv(int32_t n, const double* X, const T* b, T* x, const DevState* st,
                 hipStream_t s)
{
    hipLaunchKernelGGL(k_gemv<T>, gg(n), dim3(GB), 0, s, n, X, b, x, st);
}

template<class T>
void launch_axpy_gamg(int32_t n, T* y, const T* x, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy_gamg<T>, gg(n), dim3(GB), 0, s, n, y, x, st);
}

template<class T>
void launch_sub_gamg(int32_t n, T* y, const T* x, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_gamg<T>, gg(n), dim3(GB), 0, s, n, y, x, st);
}

template<class T>
void launch_diff_gamg(int32_t n, const T* a, const T* b, T* out, const DevState* st,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_diff_gamg<T>, gg(n), dim3(GB), 0, s, n, a, b, out, st);
}

template<class T>
void launch_copy_gamg(int32_t n, const T* a, T* b, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_gamg<T>, gg(n), dim3(GB), 0, s, n, a, b, st);
}

template<class T>
void launch_zero_gamg(int32_t n, T* a, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_gamg<T>, gg(n), dim3(GB), 0, s, n, a, st);
}

void launch_to_f32_gamg(int32_t n, const double* a, float* b, const DevState* st,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_to_f32_gamg, gg(n), dim3(GB), 0, s, n, a, b, st);
}

void launch_add_f32_gamg(int32_t n, double* y, const float* x, const DevState* st,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_add_f32_gamg, gg(n), dim3(GB), 0, s, n, y, x, st);
}

void launch_coarsest_setup(const CoarsestSys& S, hipStream_t s)
{
    hipLaunchKernelGGL(k_coarsest_setup, dim3(1), dim3(CW), 0, s, S);
}

// LDS holds the 13 natural vectors when they fit in 64 KiB
template<class T>
void launch_coarsest_krylov(const CoarsestSys& S, const T* srcLib, T* corrLib,
                            const DevState* st, hipStream_t s)
{
    const size_t bytes = 13*sizeof(double)*(size_t)S.n;
    const int useLds = bytes <= 65536 ? 1 : 0;
    hipLaunchKernelGGL(k_coarsest_krylov<T>, dim3(1), dim3(CW), useLds ? bytes : 0, s, S,
                       srcLib, corrLib, st, useLds);
}

template<class T>
void launch_lu_backsub(int32_t n, const double* LU, const int32_t* pivot, const int32_t* perm,
                       double* work, const T* srcLib, T* corrLib, const DevState* st,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_lu_backsub<T>, dim3(1), dim3(CW), 0, s, n, LU, pivot, perm, work,
                       srcLib, corrLib, st);
}

#define GI(T)                                                                            \
    template void launch_restrict<T>(int32_t, const int32_t*, const int32_t*, const T*, T*, \
                                     const DevState*, hipStream_t);                      \
    template void launch_prolong<T>(int32_t, const int32_t*, const T*, T*, const DevState*, \
                                    hipStream_t);                                        \
    template void launch_gamg_scale<T>(const RowsView&, const T*, const T*, T*, T*, const T*, \
                                       double*, DevState*, hipStream_t);                 \
    template void launch_gemv<T>(int32_t, const double*, const T*, T*, const DevState*,   \
                                 hipStream_t);                                           \
    template void launch_axpy_gamg<T>(int32_t, T*, const T*, const DevState*, hipStream_t); \
    template void launch_sub_gamg<T>(int32_t, T*, const T*, const DevState*, hipStream_t); \
    template void launch_diff_gamg<T>(int32_t, const T*, const T*, T*, const DevState*,   \
                                      hipStream_t);                                      \
    template void launch_copy_gamg<T>(int32_t, const T*, T*, const DevState*, hipStream_t); \
    template void launch_zero_gamg<T>(int32_t, T*, const DevState*, hipStream_t);         \
    template void launch_coarsest_krylov<T>(const CoarsestSys&, const T*, T*, const DevState*, \
                                            hipStream_t);                                \
    template void launch_lu_backsub<T>(int32_t, const double*, const int32_t*, const int32_t*, \
                                       double*, const T*, T*, const DevState*, hipStream_t);
GI(double)
GI(float)

} // namespace fh
