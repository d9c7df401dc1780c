// normFactor of device arrays with no allocation: the variant of
// fh_norm_factor for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// fh_norm_factor permutes psi and source into library vectors, forms A psi
// and sumA as vectors and then sums |A psi - xRef sumA| + |source - xRef
// sumA| (lduMatrixSolver.C:174-197).  Here each row gathers psi and source
// through the permutation and forms its A psi and sumA on the fly, so the
// only device memory beyond the matrix, the inputs and the output is the
// reduction state: a DevState and two partials per block, held in the
// file-scope array `scratch` (one copy per device, NOALLOC_SCRATCH_BYTES).
// The per-block values and the finalize are those of fh_norm_factor, so the
// result is the same bits.
//
// All calls on one device share that device's array: calls that overlap on
// one device (on different streams or threads) race.
#include "fh_internal.h"
#include "fh_matrix.h"
#include "fh_kernels.h"
#include "fh_solver_kernels.h"
#include "fh_reduce_dev.h"

#include <mutex>

namespace fh
{

// largest mesh: 2^24 cells; its 65536 blocks leave 2 * 65536 partials
constexpr int32_t NOALLOC_MAX_CELLS = 1 << 24;
constexpr size_t NOALLOC_STATE_BYTES = 1024;      // DevState, rounded up
constexpr size_t NOALLOC_PARTIALS = 2*(size_t)((NOALLOC_MAX_CELLS + FH_BLOCK - 1)/FH_BLOCK) + 2;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_STATE_BYTES + NOALLOC_PARTIALS*sizeof(double);
static_assert(sizeof(DevState) <= NOALLOC_STATE_BYTES, "DevState outgrew its scratch slot");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// per-device address of `scratch` (hipGetSymbolAddress with the device
// current, once per device)
static std::mutex g_scratchMutex;
static void* g_scratch[64] = {};

static void* scratch_ptr(int device)
{
    if (device < 0 || device >= 64) fail(FH_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratch[device])
    {
        void* p = nullptr;
        FH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(FH_ERR_HIP, "no device address for the scratch array");
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

// the DevState slot cleared by a kernel (no memset node for graph capture)
__global__ void k_noalloc_clear(unsigned char* state)
{
    const int i = threadIdx.x;
    if (i < (int)NOALLOC_STATE_BYTES) state[i] = 0;
}

// sum x and |x| of the library-ordered psi (psi[perm[r]]): the partials of
// k_sum_abs over the permuted vector
__global__ void k_noalloc_xref(int32_t n, const int32_t* __restrict__ perm,
                               const double* __restrict__ psi, double* partials)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s = 0.0, m = 0.0;
    if (i < n)
    {
        const double x = psi[perm[i]];
        s = x;
        m = fabs(x);
    }
    write_partials(s, m, partials, gridDim.x);
}

// per library row: A psi (k_amul's order, interfaces included), sumA
// (k_sum_a's order), then k_init_residual's two terms
__global__ void k_noalloc_rows(RowsView R, const int32_t* __restrict__ perm,
                               const double* __restrict__ diag, const double* __restrict__ val,
                               const double* __restrict__ psi,
                               const double* __restrict__ source, double* partials,
                               const DevState* st)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    double nf = 0.0, rm = 0.0;
    if (r < R.n)
    {
        const int32_t nEnt = R.rowInfo[r] >> 16;
        const int32_t base = sell_base2(R.sliceStart, r);
        double acc = diag[r]*psi[perm[r]];
        double sa = diag[r];
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            acc += val[e]*psi[perm[R.col[e]]];
            sa += val[e];
        }
        if (R.ifStart)
            for (int32_t e = R.ifStart[r]; e < R.ifStart[r + 1]; e++)
            {
                acc += val[R.ifBase + e]*psi[perm[R.ifNbr[e]]];
                sa += val[R.ifBase + e];
            }
        const double t = sa*st->sum0;
        const double b = source[perm[r]];
        nf = fabs(acc - t) + fabs(b - t);
        rm = fabs(b - acc);
    }
    write_partials(nf, rm, partials, gridDim.x);
}

__global__ void k_noalloc_out(const DevState* st, double* out)
{
    if (threadIdx.x == 0) *out = st->normFactor;
}

} // namespace fh

using namespace fh;

extern "C" int fh_norm_factor_device_noalloc(fh_matrix* M, const double* d_psi,
                                             const double* d_source, double* d_normFactor)
{
    struct Ctx { fh_matrix* M; const double* psi; const double* src; double* out; } c{
        M, d_psi, d_source, d_normFactor};
    return guarded_call([](void* v) {
        Ctx& c = *static_cast<Ctx*>(v);
        fh_matrix* M = c.M;
        if (!M || !M->mesh) fail(FH_ERR_INVALID, "NULL matrix handle");
        if (!M->hasCoeffs) fail(FH_ERR_INVALID, "matrix coefficients not set");
        if (!c.psi || !c.src || !c.out) fail(FH_ERR_INVALID, "NULL argument");
        if (M->mesh->nIfFaces > 0 && !M->ifSet)
            fail(FH_ERR_INVALID, "the mesh has coupled interfaces: set the matrix's interface "
                                 "coefficients first");
        const int32_t N = M->mesh->nCells;
        if (N > NOALLOC_MAX_CELLS)
            fail(FH_ERR_INVALID, "fh_norm_factor_device_noalloc: %d cells exceed the scratch "
                                 "maximum of %d", N, NOALLOC_MAX_CELLS);
        FH_HIP(hipSetDevice(M->mesh->device));
        unsigned char* base = static_cast<unsigned char*>(scratch_ptr(M->mesh->device));
        DevState* st = reinterpret_cast<DevState*>(base);
        double* part = reinterpret_cast<double*>(base + NOALLOC_STATE_BYTES);
        hipStream_t s = M->stream;
        const int32_t nb = reduction_blocks(N);
        hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(NOALLOC_STATE_BYTES), 0, s, base);
        hipLaunchKernelGGL(k_noalloc_xref, dim3(nb), dim3(FH_BLOCK), 0, s, N,
                           M->mesh->d_perm.p, c.psi, part);
        launch_finalize(N, part, FIN_XREF, st, 0, s);
        hipLaunchKernelGGL(k_noalloc_rows, dim3(nb), dim3(FH_BLOCK), 0, s, M->rows(),
                           M->mesh->d_perm.p, M->d_diag.p, M->d_val.p, c.psi, c.src, part, st);
        launch_finalize(N, part, FIN_INIT, st, 0, s);
        hipLaunchKernelGGL(k_noalloc_out, dim3(1), dim3(64), 0, s, st, c.out);
        FH_HIP(hipGetLastError());
    }, &c);
}
