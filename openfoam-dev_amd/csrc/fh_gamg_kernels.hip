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
__global__ void k_two_minus(int32_t n, double* __restrict__ Tm, const NsState* ns)
{
    if (ns && ns->done) return;
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    if (i >= (int64_t)n*n) return;
    const int64_t r = i % n, c = i/n;
    Tm[i] = (r == c ? 2.0 : 0.0) - Tm[i];
}

__global__ void k_ident_err(int32_t n, const double* __restrict__ AX, double* partMax,
                            const NsState* ns)
{
    if (ns && ns->done) return;
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

// X0 = A^T / (|A|_1 |A|_inf), the scale from the device state
__global__ void k_scaled_transpose(int32_t n, const double* __restrict__ A,
                                   const NsState* __restrict__ ns, double* __restrict__ X)
{
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    if (i >= (int64_t)n*n) return;
    const int64_t r = i % n, c = i/n;
    X[i] = ns->scale*A[r*n + c];
}

// column sums of |a| (one workgroup per column) and row sums (one thread
// per row) of the column-major n x n matrix, into sums[0..n) and [n..2n)
__global__ __launch_bounds__(256)
void k_ns_colsum(int32_t n, const double* __restrict__ A, double* __restrict__ sums)
{
    __shared__ double sm[256];
    const int32_t c = blockIdx.x;
    double v = 0.0;
    for (int32_t i = threadIdx.x; i < n; i += 256) v += fabs(A[(int64_t)c*n + i]);
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[c] = sm[0];
}

__global__ void k_ns_rowsum(int32_t n, const double* __restrict__ A, double* __restrict__ sums)
{
    const int32_t r = blockIdx.x*GB + threadIdx.x;
    if (r >= n) return;
    double v = 0.0;
    for (int32_t c = 0; c < n; c++) v += fabs(A[(int64_t)c*n + r]);
    sums[n + r] = v;
}

// |A|_1 (largest column sum) and |A|_inf (largest row sum): the scale of
// X0; initialises the Newton-Schulz state
__global__ __launch_bounds__(1024)
void k_ns_init(int32_t n, const double* __restrict__ sums, NsState* ns)
{
    __shared__ double sc[1024], sr[1024];
    double mc = 0.0, mr = 0.0;
    for (int32_t j = threadIdx.x; j < n; j += 1024)
    {
        mc = fmax(mc, sums[j]);
        mr = fmax(mr, sums[n + j]);
    }
    sc[threadIdx.x] = mc;
    sr[threadIdx.x] = mr;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h)
        {
            sc[threadIdx.x] = fmax(sc[threadIdx.x], sc[threadIdx.x + h]);
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        const bool ok = sc[0] > 0.0 && sr[0] > 0.0;
        ns->scale = ok ? 1.0/(sc[0]*sr[0]) : 0.0;
        ns->prev = 1e300;
        ns->err = 1e300;
        ns->done = ok ? 0 : 1;
        ns->bad = ok ? 0 : 1;
        ns->cur = 0;
        ns->its = 0;
    }
}

// the convergence test of an iteration from the partial maxima of |I - AX|
// (one workgroup; a maximum does not depend on the order): stop below
// 1e-13, or below 1e-8 once the error no longer halves; a non-finite error
// stops with bad set
__global__ __launch_bounds__(1024)
void k_ns_check(int64_t nb, const double* __restrict__ partMax, NsState* ns)
{
    if (ns->done) return;
    __shared__ double sm[1024];
    __shared__ int nonFinite;
    if (threadIdx.x == 0) nonFinite = 0;
    __syncthreads();
    double m = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 1024)
    {
        const double v = partMax[b];
        if (!isfinite(v)) nonFinite = 1;
        m = fmax(m, v);
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double err = nonFinite ? NAN : sm[0];
    ns->err = err;
    ns->its++;
    if (!isfinite(err)) { ns->bad = 1; ns->done = 1; return; }
    if (err < 1e-13 || (err < 1e-8 && err > 0.5*ns->prev)) { ns->done = 1; return; }
    ns->prev = err;
}

// after X_{k+1} = X_k T went to the other buffer: it becomes the current one
__global__ void k_ns_flip(NsState* ns)
{
    if (!ns->done && threadIdx.x == 0) ns->cur = 1 - ns->cur;
}

// the result into buffer 0
__global__ void k_ns_final(int32_t n, const double* __restrict__ X1, double* __restrict__ X0,
                           const NsState* __restrict__ ns)
{
    if (ns->cur == 0) return;
    const int64_t i = (int64_t)blockIdx.x*GB + threadIdx.x;
    if (i < (int64_t)n*n) X0[i] = X1[i];
}

// C = A B for column-major n x n FP64 matrices on the matrix cores
// (v_mfma_f64_16x16x4f64).  A workgroup of four wavefronts owns a 64 x 64
// block of C, a wavefront 32 x 32 (four 16 x 16 MFMA blocks); the inner
// dimension is staged through LDS in chunks of 16, the next chunk's loads
// issued before the current chunk's products.  The MFMA (lane l holds its
// A operand's (l%16, l/16) entry and its B operand's (l/16, l%16) entry;
// element v of its result is entry (l/16 + 4v, l%16)) computes the
// transposed product, its A operand from B and its B operand from A: lane
// l, element v then holds C(i0 + l%16, j0 + l/16 + 4v), so a store writes
// 16 consecutive rows of a column.  Skipped when the Newton-Schulz state is
// done; a null A or B is the state's current X, and outOther writes the
// other X buffer instead of C.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
#define DG_T 64
#define DG_K 16
__global__ __launch_bounds__(256)
void k_dense_gemm(int32_t n, const double* __restrict__ A, const double* __restrict__ B,
                  const double* __restrict__ X0, const double* __restrict__ X1,
                  double* __restrict__ C, int32_t outOther, const NsState* __restrict__ ns)
{
    if (ns && ns->done) return;
    const double* Xc = ns && ns->cur ? X1 : X0;
    if (!A) A = Xc;
    if (!B) B = Xc;
    if (outOther) C = const_cast<double*>(ns->cur ? X0 : X1);
    __shared__ double As[DG_K][DG_T + 1];        // A(i0 + i, k0 + k) at [k][i]
    __shared__ double Bs[DG_T][DG_K + 1];        // B(k0 + k, j0 + j) at [j][k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t i0 = (int64_t)blockIdx.x*DG_T, j0 = (int64_t)blockIdx.y*DG_T;
    const int wi = 32*(w & 1), wj = 32*(w >> 1);
    mfma_d4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[r][c] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    double av[4], bv[4];
    auto fetch = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            const int e = t + 256*u;
            const int ia = e & 63, ka = e >> 6;              // A tile: 64 rows x 16
            const int64_t gi = i0 + ia, gk = k0 + ka;
            av[u] = (gi < n && gk < n) ? A[gk*n + gi] : 0.0;
            const int kb = e & 15, jb = e >> 4;              // B tile: 16 x 64 columns
            const int64_t gk2 = k0 + kb, gj = j0 + jb;
            bv[u] = (gk2 < n && gj < n) ? B[gj*n + gk2] : 0.0;
        }
    };
    fetch(0);
    for (int64_t k0 = 0; k0 < n; k0 += DG_K)
    {
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            const int e = t + 256*u;
            As[e >> 6][e & 63] = av[u];
            Bs[e >> 4][e & 15] = bv[u];
        }
        __syncthreads();
        if (k0 + DG_K < n) fetch(k0 + DG_K);
#pragma unroll
        for (int kk = 0; kk < DG_K; kk += 4)
        {
            const int k = kk + (lane >> 4);
            double bop[2], aop[2];
#pragma unroll
            for (int r = 0; r < 2; r++) bop[r] = As[k][wi + 16*r + (lane & 15)];
#pragma unroll
            for (int c = 0; c < 2; c++) aop[c] = Bs[wj + 16*c + (lane & 15)][k];
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int c = 0; c < 2; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[c], bop[r], acc[r][c],
                                                                    0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int64_t gi = i0 + wi + 16*r + (lane & 15);
                const int64_t gj = j0 + wj + 16*c + (lane >> 4) + 4*v;
                if (gi < n && gj < n) C[gj*n + gi] = acc[r][c][v];
            }
}

// Crout decomposition with implicit partial pivoting in one workgroup, in
// the loop order of scalarMatrices.C:77-180 (every sum in the same order,
// the pivot the last row of the largest scaled magnitude): the same bits as
// that routine.  The dense natural-order matrix (row-major, in LDS) is
// gathered from the column-major library-order one; LU row-major and the
// pivots out.  A zero row sets *bad.  INLDS: the matrix in LDS (when the
// device allows n*n doubles per workgroup), else in LU itself.
template<bool INLDS>
__global__ __launch_bounds__(128)
void k_crout_lu(int32_t n, const double* __restrict__ Acm, const int32_t* __restrict__ perm,
                double* __restrict__ LU, int32_t* __restrict__ piv, int32_t* bad)
{
    extern __shared__ double Mlds[];             // n*n
    double* M = INLDS ? Mlds : LU;
    __shared__ double sc[128];
    __shared__ int32_t sMax;
    const int t = threadIdx.x;
    for (int64_t e = t; e < (int64_t)n*n; e += blockDim.x)
    {
        const int32_t c = (int32_t)(e / n), r = (int32_t)(e % n);   // library (r, c)
        M[(int64_t)perm[r]*n + perm[c]] = Acm[e];
    }
    __syncthreads();
    if (t < n)
    {
        double big = 0.0;
        for (int32_t j = 0; j < n; j++) big = fmax(big, fabs(M[(int64_t)t*n + j]));
        if (big == 0.0) *bad = 1;
        sc[t] = 1.0/big;
    }
    __syncthreads();
    for (int32_t j = 0; j < n; j++)
    {
        if (t == 0)
            for (int32_t i = 0; i < j; i++)
            {
                double sum = M[(int64_t)i*n + j];
                for (int32_t k = 0; k < i; k++) sum -= M[(int64_t)i*n + k]*M[(int64_t)k*n + j];
                M[(int64_t)i*n + j] = sum;
            }
        __syncthreads();
        if (t >= j && t < n)
        {
            double sum = M[(int64_t)t*n + j];
            for (int32_t k = 0; k < j; k++) sum -= M[(int64_t)t*n + k]*M[(int64_t)k*n + j];
            M[(int64_t)t*n + j] = sum;
        }
        __syncthreads();
        if (t == 0)
        {
            int32_t iMax = 0;
            double best = 0.0;
            for (int32_t i = j; i < n; i++)
            {
                const double s2 = sc[i]*fabs(M[(int64_t)i*n + j]);
                if (s2 >= best) { best = s2; iMax = i; }
            }
            piv[j] = iMax;
            sMax = iMax;
        }
        __syncthreads();
        const int32_t iMax = sMax;
        if (iMax != j)
        {
            for (int32_t k = t; k < n; k += blockDim.x)
            {
                const double a = M[(int64_t)j*n + k];
                M[(int64_t)j*n + k] = M[(int64_t)iMax*n + k];
                M[(int64_t)iMax*n + k] = a;
            }
            if (t == 0) sc[iMax] = sc[j];
        }
        __syncthreads();
        if (t == 0 && M[(int64_t)j*n + j] == 0.0) M[(int64_t)j*n + j] = 2.220446049250313e-16;
        __syncthreads();
        if (j != n - 1)
        {
            const double r = 1.0/M[(int64_t)j*n + j];
            if (t > j && t < n) M[(int64_t)t*n + j] *= r;
        }
        __syncthreads();
    }
    if constexpr (INLDS)
        for (int64_t e = t; e < (int64_t)n*n; e += blockDim.x) LU[e] = M[e];
}

// coarsest correction x = X b: the columns in GEMV_SPLIT fixed ranges, one
// thread per row and range summing its range in ascending order
// (coalesced over the rows of a column), then the ranges summed in order
#define GEMV_SPLIT 16
template<class T>
__global__ __launch_bounds__(64)
void k_gemv_part(int32_t n, const double* __restrict__ X, const T* __restrict__ b,
                 double* __restrict__ part, const DevState* st)
{
    if (st && st->done) return;
    const int32_t r = blockIdx.x*64 + threadIdx.x;
    if (r >= n) return;
    const int32_t cs = (n + GEMV_SPLIT - 1)/GEMV_SPLIT;
    const int32_t c0 = blockIdx.y*cs, c1 = min(n, c0 + cs);
    double s = 0.0;
    for (int32_t c = c0; c < c1; c++) s += X[(int64_t)c*n + r]*(double)b[c];
    part[(int64_t)blockIdx.y*n + r] = s;
}

template<class T>
__global__ void k_gemv_sum(int32_t n, const double* __restrict__ part, T* __restrict__ x,
                           const DevState* st)
{
    if (st && st->done) return;
    const int32_t r = blockIdx.x*GB + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int k = 0; k < GEMV_SPLIT; k++) s += part[(int64_t)k*n + r];
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

void launch_ns_init(int32_t n, const double* A, double* X0, double* sums, NsState* ns,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_ns_colsum, dim3((unsigned)n), dim3(256), 0, s, n, A, sums);
    hipLaunchKernelGGL(k_ns_rowsum, gg(n), dim3(GB), 0, s, n, A, sums);
    hipLaunchKernelGGL(k_ns_init, dim3(1), dim3(1024), 0, s, n, sums, ns);
    hipLaunchKernelGGL(k_scaled_transpose, gg((int64_t)n*n), dim3(GB), 0, s, n, A, ns, X0);
}

void launch_ns_iterations(int32_t n, const double* A, double* X0, double* X1, double* T_,
                          double* partMax, NsState* ns, int iterations, hipStream_t s)
{
    const int64_t nb = ((int64_t)n*n + GB - 1)/GB;
    const dim3 grid((unsigned)((n + DG_T - 1)/DG_T), (unsigned)((n + DG_T - 1)/DG_T));
    for (int it = 0; it < iterations; it++)
    {
        // T = A X; the test on |I - T|; T = 2I - T; X' = X T
        hipLaunchKernelGGL(k_dense_gemm, grid, dim3(256), 0, s, n, A, nullptr, X0, X1, T_, 0,
                           ns);
        hipLaunchKernelGGL(k_ident_err, dim3((unsigned)nb), dim3(GB), 0, s, n, T_, partMax, ns);
        hipLaunchKernelGGL(k_ns_check, dim3(1), dim3(1024), 0, s, nb, partMax, ns);
        hipLaunchKernelGGL(k_two_minus, gg((int64_t)n*n), dim3(GB), 0, s, n, T_, ns);
        hipLaunchKernelGGL(k_dense_gemm, grid, dim3(256), 0, s, n, nullptr, T_, X0, X1,
                           nullptr, 1, ns);
        hipLaunchKernelGGL(k_ns_flip, dim3(1), dim3(64), 0, s, ns);
    }
}

void launch_ns_final(int32_t n, const double* X1, double* X0, const NsState* ns, hipStream_t s)
{
    hipLaunchKernelGGL(k_ns_final, gg((int64_t)n*n), dim3(GB), 0, s, n, X1, X0, ns);
}

void launch_crout_lu(int32_t n, const double* Acm, const int32_t* perm, double* LU,
                     int32_t* piv, int32_t* bad, hipStream_t s)
{
    if (n > 128) fail(FH_ERR_INVALID, "internal: the one-workgroup Crout LU takes n <= 128");
    static const size_t ldsMax = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return (size_t)0;
        return (size_t)v;
    }();
    const size_t lds = sizeof(double)*(size_t)n*n;
    if (lds + 2048 <= ldsMax)
        hipLaunchKernelGGL(k_crout_lu<true>, dim3(1), dim3(128), lds, s, n, Acm, perm, LU, piv,
                           bad);
    else
        hipLaunchKernelGGL(k_crout_lu<false>, dim3(1), dim3(128), 0, s, n, Acm, perm, LU, piv,
                           bad);
}

template<class T>
void launch_gemv(int32_t n, const double* X, const T* b, T* x, double* part,
                 const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_gemv_part<T>, dim3((unsigned)((n + 63)/64), GEMV_SPLIT), dim3(64), 0,
                       s, n, X, b, part, st);
    hipLaunchKernelGGL(k_gemv_sum<T>, gg(n), dim3(GB), 0, s, n, part, x, st);
}

int64_t gemv_work_doubles(int32_t n) { return (int64_t)GEMV_SPLIT*n; }

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
    template void launch_gemv<T>(int32_t, const double*, const T*, T*, double*,           \
                                 const DevState*, hipStream_t);                          \
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
