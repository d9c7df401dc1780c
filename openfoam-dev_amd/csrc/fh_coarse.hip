// Coarse correction of the tile-local preconditioner (mixed PCG on boxes).
//
// The tile-local DIC drops the couplings between 16^3 tiles; its iteration
// count grows with the number of tiles per direction.  The two-level form
// adds a correction on the space of piecewise-constant tile functions P:
//     M^-1 r = M_tile^-1 r + P Ac^-1 P^T r,   Ac = P^T A P,
// symmetric positive (negative) definite like A, so PCG applies unchanged.
// Ac (one row per tile: the tile's diagonal sum plus twice its inner
// couplings, and the couplings to its six neighbour tiles) is assembled on
// the device in a fixed summation order, factored and inverted once per
// coefficient update (rocSOLVER Cholesky, FP64), and kept as a dense FP32
// inverse; each apply is one matrix-vector product over it.
#include "fh_kernels.h"
#include "fh_solver_kernels.h"
#include "fh_finalize_dev.h"
#include "fh_coarse.h"

namespace fh
{

#define CO_B 256

// one block per 16^3 tile: the tile's diagonal sum and inner couplings
// (counted twice), and its couplings across its lower x, y and z faces;
// each thread takes 16 cells of the tile in a fixed order, then a fixed
// block reduction
__global__ __launch_bounds__(CO_B)
void k_coarse_tiles(int32_t nx, int32_t ny, int32_t nz, int32_t nTI, int32_t nTJ,
                    const double* __restrict__ diag, const double* __restrict__ lK,
                    const double* __restrict__ lJ, const double* __restrict__ lI,
                    double* __restrict__ ctile)
{
    __shared__ double red[4][CO_B/64];
    const int32_t t = blockIdx.x;
    const int32_t TI = t % nTI, TJ = (t/nTI) % nTJ, TK = t/(nTI*nTJ);
    const int64_t sy = nx, sz = (int64_t)nx*ny;
    double dsum = 0.0, inner = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int m = 0; m < 16; m++)
    {
        const int32_t loc = threadIdx.x + CO_B*m;          // 0 .. 4095
        const int32_t x = TI*16 + (loc & 15), y = TJ*16 + ((loc >> 4) & 15);
        const int32_t z = TK*16 + (loc >> 8);
        if (x >= nx || y >= ny || z >= nz) continue;
        const int64_t c = x + sy*y + sz*z;
        dsum += diag[c];
        if (x > 0) { if (x & 15) inner += lI[c]; else cx += lI[c]; }
        if (y > 0) { if (y & 15) inner += lJ[c]; else cy += lJ[c]; }
        if (z > 0) { if (z & 15) inner += lK[c]; else cz += lK[c]; }
    }
    double v[4] = {dsum + 2.0*inner, cx, cy, cz};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
        double a = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) red[q][wid] = a;
    }
    __syncthreads();
    if (threadIdx.x < 4)
    {
        double a = 0.0;
        for (int w = 0; w < CO_B/64; w++) a += red[threadIdx.x][w];
        ctile[4*(int64_t)t + threadIdx.x] = a;
    }
}

// the sign that makes the coarse matrix positive definite: that of its
// first diagonal entry (the Laplacian is negative definite)
__device__ __forceinline__ double coarse_sign(const double* ctile)
{
    return ctile[0] < 0.0 ? -1.0 : 1.0;
}

// the dense coarse matrix times coarse_sign, column-major
__global__ void k_coarse_dense(int32_t nT, int32_t nTI, int32_t nTJ, const double* __restrict__ ctile,
                               double* __restrict__ dense)
{
    const int64_t idx = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nT*nT) return;
    const double sign = coarse_sign(ctile);
    const int32_t i = (int32_t)(idx % nT), j = (int32_t)(idx/nT);
    double v = 0.0;
    const int32_t lo = i < j ? i : j, hi = i < j ? j : i, d = hi - lo;
    const int32_t TIh = hi % nTI, TJh = (hi/nTI) % nTJ;
    if (d == 0) v = ctile[4*(int64_t)i];
    else if (d == 1 && TIh > 0) v = ctile[4*(int64_t)hi + 1];
    else if (d == nTI && TJh > 0) v = ctile[4*(int64_t)hi + 2];
    else if (d == nTI*nTJ) v = ctile[4*(int64_t)hi + 3];
    dense[idx] = sign*v;
}

// the FP32 inverse of Ac from the lower triangle of inv(sign*Ac): full
// symmetric, rows of ld >= nT floats (ld a multiple of 4, zero padding)
__global__ void k_coarse_inv32(int32_t nT, int32_t ld, const double* __restrict__ dense,
                               const double* __restrict__ ctile, float* __restrict__ inv)
{
    const int64_t idx = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nT*ld) return;
    const double sign = coarse_sign(ctile);
    const int32_t j = (int32_t)(idx % ld), i = (int32_t)(idx/ld);
    float v = 0.f;
    if (j < nT)
        v = (float)(sign*(i >= j ? dense[i + (int64_t)nT*j] : dense[j + (int64_t)nT*i]));
    inv[idx] = v;
}

void launch_coarse_tiles(int32_t nx, int32_t ny, int32_t nz, int32_t nTI, int32_t nTJ,
                         int32_t nT, const double* diag, const double* const dir[3],
                         double* ctile, hipStream_t s)
{
    hipLaunchKernelGGL(k_coarse_tiles, dim3(nT), dim3(CO_B), 0, s, nx, ny, nz, nTI, nTJ, diag,
                       dir[0], dir[1], dir[2], ctile);
}

void launch_coarse_dense(int32_t nT, int32_t nTI, int32_t nTJ, const double* ctile,
                         double* dense, hipStream_t s)
{
    const int64_t n = (int64_t)nT*nT;
    hipLaunchKernelGGL(k_coarse_dense, dim3((unsigned)((n + CO_B - 1)/CO_B)), dim3(CO_B), 0, s,
                       nT, nTI, nTJ, ctile, dense);
}

void launch_coarse_inv32(int32_t nT, int32_t ld, const double* dense, const double* ctile,
                         float* inv, hipStream_t s)
{
    const int64_t n = (int64_t)nT*ld;
    hipLaunchKernelGGL(k_coarse_inv32, dim3((unsigned)((n + CO_B - 1)/CO_B)), dim3(CO_B), 0, s,
                       nT, ld, dense, ctile, inv);
}

// ---------------------------------------------------------------------------
// Block-tridiagonal inverse of s*Ac (no vendor library).  In natural tile
// order Ac is block tridiagonal over the K = nTK planes of P = nTI*nTJ
// tiles: diagonal blocks A_k (five-point couplings within the plane) and,
// between planes k-1 and k, the diagonal matrix C_k of the z couplings
// (ctile[4t + 3] of the tiles of plane k).  With the Schur complements
//     S_0 = A_0,  S_k = A_k - C_k S_{k-1}^-1 C_k,  U_k = S_k^-1,
// the inverse Z = (s*Ac)^-1 follows by the backward block recurrence
//     Z_{K-1,K-1} = U_{K-1},
//     Z_{j,k} = -Z_{j,k+1} C_{k+1} U_k          (j > k),
//     Z_{k,k} = U_k - (C_{k+1} U_k)^T Z_{k+1,k},
// of which only the blocks j >= k are formed.  The K inversions of P x P
// SPD matrices run one after another; each is split 2 x 2 at 128 rows
// (S = [A B; B^T D]: A^-1 and T^-1 for T = D - B^T A^-1 B by Gauss-Jordan
// in the registers of one workgroup each, joined by block products), all in
// FP64; the recurrence is batches of FP64 block products.
// ---------------------------------------------------------------------------

// S_k = s*A_k - C_k U_{k-1} C_k (FP64), one entry per thread
__global__ void k_coarse_schur_form(int32_t k, int32_t P, int32_t nTI,
                                    const double* __restrict__ ctile,
                                    const double* __restrict__ U, double* __restrict__ S)
{
    const int64_t e = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)P*P) return;
    const int i = (int)(e / P), j = (int)(e % P);
    const double sign = coarse_sign(ctile);
    const double* ct = ctile + 4*(int64_t)k*P;
    const int lo = min(i, j), hi = max(i, j), d = hi - lo;
    double v = 0.0;
    if (d == 0) v = ct[4*i];
    else if (d == 1 && hi % nTI > 0) v = ct[4*hi + 1];
    else if (d == nTI) v = ct[4*hi + 2];
    v *= sign;
    if (k > 0)
    {
        const double* Up = U + (int64_t)(k - 1)*P*P;
        v -= (sign*ct[4*i + 3])*Up[e]*(sign*ct[4*j + 3]);
    }
    S[e] = v;
}

// Gauss-Jordan inverse of an n x n SPD block (n <= H, rows of ld doubles)
// without pivoting, in FP64 in the registers of one workgroup of
// (H/4)*(H/16) threads.  A non-positive pivot sets *bad (the caller falls
// back to the one-level preconditioner).
//
// Thread (tr, tc) = (t / (H/16), t % (H/16)) owns rows 4tr .. 4tr+3 and
// columns 16tc .. 16tc+15, so the pivot row and column it reads (and, as
// their owner, writes) are contiguous in LDS.  The pivot loop is unrolled
// over p mod 16, so the owners of the pivot row (tr = p/4, row p mod 4) and
// column (tc = p/16, column p mod 16) index their registers statically: a
// step is the rank-one update of every entry (one FMA each), then the
// owners overwrite the pivot row and column with the scaled values.
template<class T, int RPT, int CPT, int CU>
__device__ __forceinline__ bool gj_step(T (&a)[RPT][CPT], int tr, int tc, int p, T (*rowp)[256],
                                        T (*colp)[256], int32_t* bad)
{
    constexpr int RA = CU % RPT;               // the pivot row's register row
    constexpr int B = CU & 1;                  // pivot buffer: the step's parity
    const int prow = p / RPT, pcol = p / CPT;  // owners: tr == prow, tc == pcol
    if (tr == prow)
    {
#pragma unroll
        for (int c = 0; c < CPT; c++) rowp[B][CPT*tc + c] = a[RA][c];
    }
    if (tc == pcol)
    {
#pragma unroll
        for (int ra = 0; ra < RPT; ra++) colp[B][RPT*tr + ra] = a[ra][CU];
    }
    // one barrier per step: the next step writes the other buffer, and the
    // step after it this one only once every thread is past the next
    // step's barrier, i.e. done reading this step's values
    __syncthreads();
    const T piv = rowp[B][p];
    if (!(piv > T(0)))
    {
        // every thread reads the same pivot: the block leaves together
        if (threadIdx.x == 0) *bad = 1;
        return false;
    }
    const T ip = T(1)/piv;
    T cp[RPT], rp[CPT];
#pragma unroll
    for (int ra = 0; ra < RPT; ra++) cp[ra] = colp[B][RPT*tr + ra];
#pragma unroll
    for (int c = 0; c < CPT; c++) rp[c] = rowp[B][CPT*tc + c]*ip;
#pragma unroll
    for (int ra = 0; ra < RPT; ra++)
#pragma unroll
        for (int c = 0; c < CPT; c++) a[ra][c] = fma(-cp[ra], rp[c], a[ra][c]);
    // the pivot row: the old row over the pivot; the pivot column: -old
    // column over the pivot; the pivot entry: its reciprocal
    if (tr == prow)
    {
#pragma unroll
        for (int c = 0; c < CPT; c++) a[RA][c] = rp[c];
    }
    if (tc == pcol)
    {
#pragma unroll
        for (int ra = 0; ra < RPT; ra++) a[ra][CU] = -cp[ra]*ip;
        if (tr == prow) a[RA][CU] = ip;
    }
    return true;
}

// the pivots base .. base+CPT-1 (p mod CPT static: the owners index their
// registers statically; CPT is a multiple of RPT)
template<class T, int RPT, int CPT, int CU>
__device__ __forceinline__ bool gj_unrolled(T (&a)[RPT][CPT], int tr, int tc, int base, int P,
                                            T (*rowp)[256], T (*colp)[256], int32_t* bad)
{
    if constexpr (CU < CPT)
    {
        const int p = base + CU;
        if (p >= P) return true;
        if (!gj_step<T, RPT, CPT, CU>(a, tr, tc, p, rowp, colp, bad)) return false;
        return gj_unrolled<T, RPT, CPT, CU + 1>(a, tr, tc, base, P, rowp, colp, bad);
    }
    return true;
}

// one workgroup of (H/RPT)*(H/CPT) threads, thread (tr, tc) owning rows
// RPT tr .. RPT tr + RPT-1 and columns CPT tc .. CPT tc + CPT-1
template<int H, int RPT, int CPT>
__global__ __launch_bounds__((H/RPT)*(H/CPT))
void k_coarse_gj(int32_t n, const double* __restrict__ src, int64_t lds, double* __restrict__ dst,
                 int64_t ldd, int32_t* bad)
{
    static_assert(H <= 256 && H % CPT == 0 && CPT % RPT == 0, "Gauss-Jordan layout");
    __shared__ __align__(16) double rowp[2][256], colp[2][256];
    constexpr int NCG = H/CPT;
    const int tr = threadIdx.x / NCG, tc = threadIdx.x % NCG;
    double a[RPT][CPT];
#pragma unroll
    for (int ra = 0; ra < RPT; ra++)
    {
        const int i = RPT*tr + ra;
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            const int j = CPT*tc + c;
            a[ra][c] = (i < n && j < n) ? src[(int64_t)i*lds + j] : 0.0;
        }
    }
#pragma unroll 1
    for (int base = 0; base < n; base += CPT)
        if (!gj_unrolled<double, RPT, CPT, 0>(a, tr, tc, base, n, rowp, colp, bad)) return;
#pragma unroll
    for (int ra = 0; ra < RPT; ra++)
    {
        const int i = RPT*tr + ra;
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            const int j = CPT*tc + c;
            if (i < n && j < n) dst[(int64_t)i*ldd + j] = a[ra][c];
        }
    }
}

// the Gauss-Jordan workgroup of one 128-block: 256 threads of 4 x 16
// entries (1024 threads of 2 x 8 measured 7% slower: the steps are bound by
// the pivot broadcast and the barrier, not by the rank-one update)
static void launch_gj128(int32_t n, const double* src, int64_t lds, double* dst, int64_t ldd,
                         int32_t* bad, hipStream_t s)
{
    hipLaunchKernelGGL((k_coarse_gj<128, 4, 16>), dim3(1), dim3(256), 0, s, n, src, lds, dst,
                       ldd, bad);
}

// the inverse of S = [A B; B^T D] (A: h x h) from A^-1, W = A^-1 B T^-1,
// T^-1 and the top-left block A^-1 + W (A^-1 B)^T:  [TL -W; -W^T T^-1]
__global__ void k_coarse_assemble(int32_t P, int32_t h, const double* __restrict__ TL,
                                  const double* __restrict__ W, const double* __restrict__ Ti,
                                  double* __restrict__ U0)
{
    const int64_t e = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)P*P) return;
    const int i = (int)(e / P), j = (int)(e % P), h2 = P - h;
    double v;
    if (i < h && j < h) v = TL[(int64_t)i*h + j];
    else if (i < h) v = -W[(int64_t)i*h2 + (j - h)];
    else if (j < h) v = -W[(int64_t)j*h2 + (i - h)];
    else v = Ti[(int64_t)(i - h)*h2 + (j - h)];
    U0[e] = v;
}

// Batched FP64 block products of the recurrence on the matrix cores
// (v_mfma_f64_16x16x4f64):  out = beta*D + alpha*op(A) diag(c) op(B)  for
// blockIdx.z = 0 .. nb-1 blocks of M x N outputs over an inner dimension Kd
// (row-major, leading dimensions ldA, ldB, ldD, ldO; block z of each
// operand at z*stride).  op(X) = X^T when transX.  out must not overlap A,
// B or D.  A workgroup of four wavefronts owns 64 x 64 outputs, a
// wavefront 32 x 32 (four 16 x 16 MFMA blocks); the inner dimension is
// staged through LDS in chunks of 64 (c applied to the B chunk as it is
// staged), the next chunk's loads issued before the current chunk's
// products.  The products here are small (128 or 256 square), so a launch
// is bound by the chain of chunk loads: chunks of 64 make it two or four
// round trips (chunks of 16 took 33.6 us per launch).  MFMA lane layout: lane l holds its A operand's (l%16, l/16)
// entry and its B operand's (l/16, l%16) entry; element v of the result is
// entry (l/16 + 4v, l%16), so a store writes 16 consecutive outputs of a row.
#define GM_T 64
#define GM_K 64
#define GM_U (GM_T*GM_K/256)   // chunk values per thread and operand
struct BlockGemm
{
    int32_t M, N, Kd;
    bool transA, transB;
    double alpha, beta;
    const double* A; int64_t ldA, strideA;
    const double* B; int64_t ldB, strideB;
    const double* c;
    const double* D; int64_t ldD, strideD;
    double* out; int64_t ldO, strideO;
};

typedef double co_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256)
void k_coarse_block_gemm(BlockGemm g)
{
    __shared__ double As[GM_K][GM_T + 1];      // op(A)(i0 + i, k0 + k) at [k][i]
    __shared__ double Bs[GM_K][GM_T + 1];      // c(k) op(B)(k0 + k, j0 + j) at [k][j]
    const int z = blockIdx.z;
    const double* Ab = g.A + z*g.strideA;
    const double* Bb = g.B + z*g.strideB;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int i0 = blockIdx.x*GM_T, j0 = blockIdx.y*GM_T;
    const int wi = 32*(w & 1), wj = 32*(w >> 1);
    co_d4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[r][c] = co_d4{0.0, 0.0, 0.0, 0.0};
    double av[GM_U], bv[GM_U];
    // the chunk's operands, GM_U values per thread each: the index that is
    // contiguous in memory runs fastest over the threads
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < GM_U; u++)
        {
            const int e = t + 256*u;
            int ia, ka;
            if (g.transA) { ia = e % GM_T; ka = e / GM_T; }     // A(k, i): i contiguous
            else { ka = e % GM_K; ia = e / GM_K; }              // A(i, k): k contiguous
            const int gi = i0 + ia, gk = k0 + ka;
            av[u] = (gi < g.M && gk < g.Kd)
                  ? (g.transA ? Ab[(int64_t)gk*g.ldA + gi] : Ab[(int64_t)gi*g.ldA + gk]) : 0.0;
            int jb, kb;
            if (g.transB) { kb = e % GM_K; jb = e / GM_K; }     // B(j, k): k contiguous
            else { jb = e % GM_T; kb = e / GM_T; }              // B(k, j): j contiguous
            const int gj = j0 + jb, gk2 = k0 + kb;
            bv[u] = (gj < g.N && gk2 < g.Kd)
                  ? g.c[gk2]*(g.transB ? Bb[(int64_t)gj*g.ldB + gk2] : Bb[(int64_t)gk2*g.ldB + gj])
                  : 0.0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < g.Kd; k0 += GM_K)
    {
#pragma unroll
        for (int u = 0; u < GM_U; u++)
        {
            const int e = t + 256*u;
            if (g.transA) As[e / GM_T][e % GM_T] = av[u];
            else As[e % GM_K][e / GM_K] = av[u];
            if (g.transB) Bs[e % GM_K][e / GM_K] = bv[u];
            else Bs[e / GM_T][e % GM_T] = bv[u];
        }
        __syncthreads();
        if (k0 + GM_K < g.Kd) fetch(k0 + GM_K);
#pragma unroll
        for (int kk = 0; kk < GM_K; kk += 4)
        {
            const int k = kk + (lane >> 4);
            double aop[2], bop[2];
#pragma unroll
            for (int r = 0; r < 2; r++) aop[r] = As[k][wi + 16*r + (lane & 15)];
#pragma unroll
            for (int c = 0; c < 2; c++) bop[c] = Bs[k][wj + 16*c + (lane & 15)];
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int c = 0; c < 2; c++)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[r], bop[c], acc[r][c],
                                                                    0, 0, 0);
        }
        __syncthreads();
    }
    double* Ob = g.out + z*g.strideO;
    const double* Db = g.D ? g.D + z*g.strideD : nullptr;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++)
            {
                const int gi = i0 + wi + 16*r + (lane >> 4) + 4*v;
                const int gj = j0 + wj + 16*c + (lane & 15);
                if (gi < g.M && gj < g.N)
                {
                    double rr = g.alpha*acc[r][c][v];
                    if (Db) rr += g.beta*Db[(int64_t)gi*g.ldD + gj];
                    Ob[(int64_t)gi*g.ldO + gj] = rr;
                }
            }
}

// cvec[t] = s * (coupling of tile t to its lower-z neighbour)
__global__ void k_coarse_cvec(int32_t nT, const double* __restrict__ ctile,
                              double* __restrict__ cvec)
{
    const int32_t t = blockIdx.x*blockDim.x + threadIdx.x;
    if (t < nT) cvec[t] = coarse_sign(ctile)*ctile[4*(int64_t)t + 3];
}

// Z_{K-1,K-1} = U_{K-1} into the dense row-major Z (n = K*P)
__global__ void k_coarse_zlast(int32_t P, int32_t K, const double* __restrict__ U,
                               double* __restrict__ Z)
{
    const int64_t e = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)P*P) return;
    const int64_t n = (int64_t)K*P;
    const int32_t a = (int32_t)(e / P), b = (int32_t)(e % P);
    Z[((int64_t)(K - 1)*P + a)*n + (int64_t)(K - 1)*P + b] = U[(int64_t)(K - 1)*P*P + e];
}

// the full FP32 inverse of Ac (sign * Z), rows of ld floats, from the
// blocks j >= k of Z: 32 x 32 output tiles; a tile above the block diagonal
// is the transpose of its mirror tile, read whole rows at a time into LDS
// and written transposed (P is a multiple of 32 or the tiles are masked)
#define CI_T 32
__global__ __launch_bounds__(256)
void k_coarse_inv32_blocks(int32_t nT, int32_t P, int32_t ld, const double* __restrict__ Z,
                           const double* __restrict__ ctile, float* __restrict__ inv)
{
    __shared__ double tileT[CI_T][CI_T + 1];
    const double sign = coarse_sign(ctile);
    const int32_t i0 = blockIdx.y*CI_T, j0 = blockIdx.x*CI_T;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;      // 32 x 8
    if (i0 >= nT) return;
    // the plane order of the tile's rows and columns (one plane per tile
    // row range when 32 divides P; else decided per element below)
    const bool direct = (i0/P) >= (j0/P) && ((i0 + CI_T - 1)/P) >= ((j0 + CI_T - 1)/P)
                        && (P % CI_T) == 0;
    if (direct || (P % CI_T) != 0)
    {
        for (int r = ty; r < CI_T; r += 8)
        {
            const int32_t i = i0 + r, j = j0 + tx;
            if (i >= nT || j >= ld) continue;
            float v = 0.f;
            if (j < nT)
                v = (float)(sign*((i/P) >= (j/P) ? Z[(int64_t)i*nT + j] : Z[(int64_t)j*nT + i]));
            inv[(int64_t)i*ld + j] = v;
        }
        return;
    }
    // transposed tile: read rows j0.. of Z at columns i0.. (coalesced)
    for (int r = ty; r < CI_T; r += 8)
    {
        const int32_t j = j0 + r, i = i0 + tx;
        tileT[r][tx] = (j < nT && i < nT) ? Z[(int64_t)j*nT + i] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < CI_T; r += 8)
    {
        const int32_t i = i0 + r, j = j0 + tx;
        if (i >= nT || j >= ld) continue;
        inv[(int64_t)i*ld + j] = j < nT ? (float)(sign*tileT[tx][r]) : 0.f;
    }
}

__global__ void k_coarse_fill(int64_t n, double v, double* __restrict__ x)
{
    const int64_t i = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

static void gemm_launch(const BlockGemm& g, unsigned nb, hipStream_t s)
{
    const dim3 grid((unsigned)((g.M + GM_T - 1)/GM_T), (unsigned)((g.N + GM_T - 1)/GM_T), nb);
    hipLaunchKernelGGL(k_coarse_block_gemm, grid, dim3(256), 0, s, g);
}

static BlockGemm gemm_sq(int32_t P)
{
    BlockGemm g{};
    g.M = g.N = g.Kd = P;
    return g;
}

// U0 = S^-1 for the P x P SPD matrix S (FP64): one Gauss-Jordan workgroup
// for P <= 128, else the 2 x 2 split at h = 128
static void coarse_spd_inverse(int32_t P, const double* S, double* U0, double* t,
                               const double* ones, int32_t* bad, hipStream_t s)
{
    if (P <= 128)
    {
        launch_gj128(P, S, (int64_t)P, U0, (int64_t)P, bad, s);
        return;
    }
    const int32_t h = 128, h2 = P - h;
    double* Ai = t;                         // h x h
    double* X = Ai + h*h;                   // h x h2: A^-1 B
    double* T = X + h*h2;                   // h2 x h2: D - B^T A^-1 B, then T^-1
    double* Ti = T + h2*h2;
    double* W = Ti + h2*h2;                 // h x h2: X T^-1
    double* TL = W + h*h2;                  // h x h
    launch_gj128(h, S, (int64_t)P, Ai, (int64_t)h, bad, s);
    BlockGemm x{};
    x.M = h; x.N = h2; x.Kd = h; x.alpha = 1.0; x.c = ones;
    x.A = Ai; x.ldA = h; x.B = S + h; x.ldB = P; x.out = X; x.ldO = h2;
    gemm_launch(x, 1, s);
    BlockGemm tt{};
    tt.M = h2; tt.N = h2; tt.Kd = h; tt.transA = true; tt.alpha = -1.0; tt.beta = 1.0; tt.c = ones;
    tt.A = S + h; tt.ldA = P; tt.B = X; tt.ldB = h2; tt.D = S + (int64_t)h*P + h; tt.ldD = P;
    tt.out = T; tt.ldO = h2;
    gemm_launch(tt, 1, s);
    launch_gj128(h2, T, (int64_t)h2, Ti, (int64_t)h2, bad, s);
    BlockGemm w{};
    w.M = h; w.N = h2; w.Kd = h2; w.alpha = 1.0; w.c = ones;
    w.A = X; w.ldA = h2; w.B = Ti; w.ldB = h2; w.out = W; w.ldO = h2;
    gemm_launch(w, 1, s);
    BlockGemm tl{};
    tl.M = h; tl.N = h; tl.Kd = h2; tl.transB = true; tl.alpha = 1.0; tl.beta = 1.0; tl.c = ones;
    tl.A = W; tl.ldA = h2; tl.B = X; tl.ldB = h2; tl.D = Ai; tl.ldD = h; tl.out = TL; tl.ldO = h;
    gemm_launch(tl, 1, s);
    hipLaunchKernelGGL(k_coarse_assemble, dim3((unsigned)((P*P + CO_B - 1)/CO_B)), dim3(CO_B), 0,
                       s, P, h, TL, W, Ti, U0);
}

bool launch_coarse_block_inverse(int32_t nT, int32_t nTI, int32_t nTJ, int32_t ld,
                                 const double* ctile, double* U, double* Z, double* work,
                                 float* inv, int32_t* bad, hipStream_t s)
{
    const int32_t P = nTI*nTJ, K = nT/P;
    if (P > 256 || K < 2 || K*P != nT) return false;
    const int64_t PP = (int64_t)P*P;
    double* S = work;               // S_k (FP64)
    double* ones = S + PP;          // P ones
    double* cvec = ones + P;        // nT: s * z couplings
    double* split = cvec + nT;      // the 2 x 2 split's blocks
    hipLaunchKernelGGL(k_coarse_fill, dim3((P + CO_B - 1)/CO_B), dim3(CO_B), 0, s, (int64_t)P,
                       1.0, ones);
    for (int32_t k = 0; k < K; k++)
    {
        hipLaunchKernelGGL(k_coarse_schur_form, dim3((unsigned)((PP + CO_B - 1)/CO_B)), dim3(CO_B),
                           0, s, k, P, nTI, ctile, U, S);
        coarse_spd_inverse(P, S, U + (int64_t)k*PP, split, ones, bad, s);
    }
    hipLaunchKernelGGL(k_coarse_cvec, dim3((nT + CO_B - 1)/CO_B), dim3(CO_B), 0, s, nT, ctile,
                       cvec);
    const int64_t n = (int64_t)K*P;
    hipLaunchKernelGGL(k_coarse_zlast, dim3((P*P + CO_B - 1)/CO_B), dim3(CO_B), 0, s, P, K, U, Z);
    for (int32_t k = K - 2; k >= 0; k--)
    {
        const double* Uk = U + (int64_t)k*PP;
        const double* c1 = cvec + (int64_t)(k + 1)*P;
        // Z_{j,k} = -Z_{j,k+1} C_{k+1} U_k for j = k+1 .. K-1
        BlockGemm a = gemm_sq(P);
        a.alpha = -1.0; a.beta = 0.0;
        a.A = Z + (int64_t)(k + 1)*P*n + (int64_t)(k + 1)*P; a.ldA = n; a.strideA = (int64_t)P*n;
        a.B = Uk; a.ldB = P; a.strideB = 0;
        a.c = c1;
        a.out = Z + (int64_t)(k + 1)*P*n + (int64_t)k*P; a.ldO = n; a.strideO = (int64_t)P*n;
        gemm_launch(a, (unsigned)(K - 1 - k), s);
        // Z_{k,k} = U_k - U_k^T C_{k+1} Z_{k+1,k}
        BlockGemm b = gemm_sq(P);
        b.transA = true; b.alpha = -1.0; b.beta = 1.0;
        b.A = Uk; b.ldA = P;
        b.B = Z + (int64_t)(k + 1)*P*n + (int64_t)k*P; b.ldB = n;
        b.c = c1;
        b.D = Uk; b.ldD = P;
        b.out = Z + (int64_t)k*P*n + (int64_t)k*P; b.ldO = n;
        gemm_launch(b, 1, s);
    }
    hipLaunchKernelGGL(k_coarse_inv32_blocks,
                       dim3((unsigned)((ld + CI_T - 1)/CI_T), (unsigned)((nT + CI_T - 1)/CI_T)),
                       dim3(256), 0, s, nT, P, ld, Z, ctile, inv);
    return true;
}

// S, ones, cvec and the six blocks of the 2 x 2 split (each at most
// 128 x 128)
int64_t coarse_block_work_doubles(int32_t nT, int32_t P)
{
    return (int64_t)P*P + P + nT + 6*128*128;
}

// z = Ainv rc (one wavefront per row, four rows per block) and sum z*rc;
// block 0 also adds the fine partials of the tile-local apply (sum w*r, nF
// of them), so that the last block finalizes the whole w.r of the
// two-level apply (op FIN_PCG_WARA)
// NIT > 0 (ld <= 256*NIT): the wavefront's row is loaded into registers,
// all NIT 16-byte loads per lane in flight, before rc is staged, so the
// row's latency overlaps the staging; NIT = 0: one load per step of a loop
// over the row (any ld).  Both sum in the same order.
template<int NIT>
__global__ __launch_bounds__(CO_B)
void k_coarse_apply(int32_t nT, int32_t ld, const float* __restrict__ inv,
                    const double* __restrict__ rc, float* __restrict__ z,
                    const double* __restrict__ fine, int32_t nF, double* partials,
                    FusedFinalize F, const DevState* st)
{
    if (st && st->done) return;
    extern __shared__ float rcs[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t row = blockIdx.x*4 + wid;
    const float* a = inv + (int64_t)row*ld;
    float4 v[NIT > 0 ? NIT : 1];
    if constexpr (NIT > 0)
    {
        if (row < nT)
        {
#pragma unroll
            for (int it = 0; it < NIT; it++)
            {
                const int32_t j = lane*4 + 256*it;
                v[it] = j < ld ? *reinterpret_cast<const float4*>(a + j)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
    }
    for (int32_t j = threadIdx.x; j < ld; j += CO_B) rcs[j] = j < nT ? (float)rc[j] : 0.f;
    __syncthreads();
    double part = 0.0;
    if (row < nT)
    {
        float acc = 0.f;
        if constexpr (NIT > 0)
        {
#pragma unroll
            for (int it = 0; it < NIT; it++)
            {
                const int32_t j = lane*4 + 256*it;
                if (j < ld)
                    acc += v[it].x*rcs[j] + v[it].y*rcs[j + 1] + v[it].z*rcs[j + 2]
                         + v[it].w*rcs[j + 3];
            }
        }
        else
        {
            for (int32_t j = lane*4; j < ld; j += 256)
            {
                v[0] = *reinterpret_cast<const float4*>(a + j);
                acc += v[0].x*rcs[j] + v[0].y*rcs[j + 1] + v[0].z*rcs[j + 2] + v[0].w*rcs[j + 3];
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0)
        {
            z[row] = acc;
            part = (double)acc*rc[row];
        }
    }
    __shared__ double red[CO_B/64];
    if (lane == 0) red[wid] = part;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < CO_B/64; w++) t += red[w];
    if (blockIdx.x == 0)
    {
        // the fine partials, fixed order
        double f = 0.0;
        for (int32_t i = threadIdx.x; i < nF; i += CO_B) f += fine[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
        __syncthreads();
        if (lane == 0) red[wid] = f;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 0; w < CO_B/64; w++) t += red[w];
    }
    partials_arrive(t, 0.0, partials, gridDim.x, blockIdx.x, F);
}

__global__ __launch_bounds__(CO_B)
void k_add_tile_correction(int32_t nx, int32_t ny, int64_t n, int32_t nTI, int32_t nTJ,
                           const float* __restrict__ z, float* __restrict__ w,
                           const DevState* st)
{
    const int64_t c = (int64_t)blockIdx.x*CO_B + threadIdx.x;
    if (c >= n || (st && st->done)) return;
    const int32_t i = (int32_t)(c % nx);
    const int64_t jk = c/nx;
    const int32_t j = (int32_t)(jk % ny), k = (int32_t)(jk/ny);
    w[c] += z[(i >> 4) + nTI*((j >> 4) + nTJ*(k >> 4))];
}

// p = (w + z) + beta p (PCG.C:137-146 on the two-level apply), w read from
// rows padded to nxp cells (the tile-local apply's padded output), z the
// coarse correction of the cell's 16^3 tile or null: the arithmetic of the
// unpadding copy, k_add_tile_correction and k_pcg_p in sequence
__global__ __launch_bounds__(CO_B)
void k_pcg_p_padded(int32_t nx, int32_t ny, int64_t n, int32_t nxp, int32_t nTI, int32_t nTJ,
                    const float* __restrict__ wp, const float* __restrict__ z,
                    float* __restrict__ p, const DevState* st)
{
    if (st->done) return;
    const int64_t c = (int64_t)blockIdx.x*CO_B + threadIdx.x;
    if (c >= n) return;
    const int32_t i = (int32_t)(c % nx);
    const int64_t row = c/nx;
    float w = wp[i + row*nxp];
    if (z)
    {
        const int32_t j = (int32_t)(row % ny), k = (int32_t)(row/ny);
        w += z[(i >> 4) + nTI*((j >> 4) + nTJ*(k >> 4))];
    }
    if (st->nIterations == 0) p[c] = w;
    else
    {
        const float beta = (float)st->beta;
        p[c] = w + beta*p[c];
    }
}

void launch_pcg_p_padded(int32_t nx, int32_t ny, int32_t nz, int32_t nxp, const float* wp,
                         const float* z, float* p, const DevState* st, hipStream_t s)
{
    const int64_t n = (int64_t)nx*ny*nz;
    if (n > 0)
        hipLaunchKernelGGL(k_pcg_p_padded, dim3((unsigned)((n + CO_B - 1)/CO_B)), dim3(CO_B), 0,
                           s, nx, ny, n, nxp, (nx + 15)/16, (ny + 15)/16, wp, z, p, st);
}

void launch_add_tile_correction(int32_t nx, int32_t ny, int32_t nz, const float* z, float* w,
                                const DevState* st, hipStream_t s)
{
    const int64_t n = (int64_t)nx*ny*nz;
    if (n > 0)
        hipLaunchKernelGGL(k_add_tile_correction, dim3((unsigned)((n + CO_B - 1)/CO_B)),
                           dim3(CO_B), 0, s, nx, ny, n, (nx + 15)/16, (ny + 15)/16, z, w, st);
}

int32_t launch_coarse_apply(int32_t nT, int32_t ld, const float* inv, const double* rc,
                            float* z, const double* fine, int32_t nF, double* partials,
                            const FusedFinalize& F, const DevState* st, hipStream_t s)
{
    const int32_t nb = (nT + 3)/4;
    const size_t sh = (size_t)ld*sizeof(float);
    if (ld <= 1024)
        hipLaunchKernelGGL(k_coarse_apply<4>, dim3(nb), dim3(CO_B), sh, s, nT, ld, inv, rc, z,
                           fine, nF, partials, F, st);
    else if (ld <= 2048)
        hipLaunchKernelGGL(k_coarse_apply<8>, dim3(nb), dim3(CO_B), sh, s, nT, ld, inv, rc, z,
                           fine, nF, partials, F, st);
    else if (ld <= 4096)
        hipLaunchKernelGGL(k_coarse_apply<16>, dim3(nb), dim3(CO_B), sh, s, nT, ld, inv, rc, z,
                           fine, nF, partials, F, st);
    else
        hipLaunchKernelGGL(k_coarse_apply<0>, dim3(nb), dim3(CO_B), sh, s, nT, ld, inv, rc, z,
                           fine, nF, partials, F, st);
    return nb;
}

} // namespace fh
