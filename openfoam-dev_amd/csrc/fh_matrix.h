// Internal definition of fh_matrix: the device-resident coefficients of one
// lduMatrix on a mesh, its cached preconditioner factors and the work space
// of the solvers that run on it.
#pragma once

#include "fh_kernels.h"
#include "fh_solver_kernels.h"
#include "fh_tile_sweep.h"

#include <algorithm>
#include <vector>

struct fh_matrix
{
    fh_mesh* mesh = nullptr;
    hipStream_t stream = nullptr;
    bool ownsStream = true;                // false: a GAMG level sharing its parent's
    void* gamg = nullptr;                  // GAMG level matrices (fh_gamg.cpp)
    hipEvent_t evStart = nullptr, evStop = nullptr;

    bool hasCoeffs = false;
    bool symmetric = true;
    bool diagonalOnly = false;

    // FP64 coefficients: face order (as handed over) and the library rows
    fh::DBuf<double> d_upperF, d_lowerF;  // d_lowerF empty when symmetric
    fh::DBuf<double> d_diag;              // [N], library numbering
    fh::DBuf<double> d_val;               // [nnzPad] SELL
    fh::DBuf<double> d_valT;              // transposed roles (Tmul, PBiCG)
    bool valTValid = false;
    fh::DBuf<double> d_sumA;
    bool sumAValid = false;
    // coupled interfaces: interfaceBouCoeffs and interfaceIntCoeffs in the
    // caller's interface-face order (set before a solve on a coupled mesh)
    fh::DBuf<double> d_ifBou, d_ifInt;
    bool ifSet = false;

    // FP32 copies for the mixed-precision inner solves
    fh::DBuf<float> d_diag32, d_val32;
    bool f32Valid = false;
    // FP32 direction-packed coefficients of hex boxes (tile-local sweeps):
    // lower K, J, I and upper I, J, K per cell
    fh::DBuf<float> d_dir32[6];
    bool dir32Valid = false;
    // FP16 lower K, J, I coefficients of a symmetric box (tile-local inner
    // preconditioner)
    fh::DBuf<__half> d_dir16[3];
    bool dir16Valid = false;
    // the same on rows padded to a multiple of 16 cells (boxes whose nx is
    // not one), with the factor, for the fused tile-local kernels: pad
    // cells hold zero coefficients and factor; x and w staged padded
    fh::DBuf<float> d_padDir32[6], d_padRD, d_padX, d_padW;
    fh::DBuf<__half> d_padDir16[3];
    bool padValid = false;
    // set by a caller that reads w padded: the apply then leaves w in
    // d_padW (padLeft) instead of copying it back
    bool padKeep = false, padLeft = false;
    // set by a caller whose kernels wrote x padded already (d_padX)
    bool padXFresh = false;
    // FP16 storage scale 2^e of those coefficients: state 0 not decided, 1
    // FP16 at halfScale, -1 FP32 (the magnitudes span more than the FP16
    // normal range holds)
    int halfState = 0;
    float halfScale = 1.f;
    // FP64 direction-packed coefficients of hex boxes (FP64 box product and
    // sweeps): lower K, J, I, and for asymmetric matrices upper I, J, K
    fh::DBuf<double> d_dir64[6];
    bool dir64Valid = false;
    // tile depth of the global preconditioner sweeps on boxes
    int32_t preBz64 = 0, preBz32 = 0;
    // coarse correction of the tile-local preconditioner (mixed PCG on
    // symmetric boxes): the coarse matrix over 16^3 tiles as a dense FP32
    // inverse (rows of cLd floats); state 0 not built, 1 ready, -1 not used
    int coarseState = 0;
    int32_t cN = 0, cLd = 0;
    fh::DBuf<float> d_cinv, d_cz;
    fh::DBuf<double> d_cdense, d_ctile, d_crc;
    fh::DBuf<double> d_cU, d_cwork;        // block-tridiagonal inverse: Schur inverses, work
    fh::DBuf<int32_t> d_cinfo;
    void* cblas = nullptr;                 // rocblas handle of the factorization

    // preconditioner / smoother factors (library numbering)
    // index: 0 DIC/FDIC, 1 DILU, 2 diagonal, 3 tile-local DIC, 4 tile-local DILU
    fh::DBuf<double> d_rD[5];
    bool rDValid[5] = {false, false, false, false, false};
    fh::DBuf<float> d_rD32[5];
    bool rD32Valid[5] = {false, false, false, false, false};

    // pipelined sweep synchronisation
    fh::DBuf<unsigned long long> d_sync;  // [0] ticket, [1..nParts] progress
    unsigned long long epoch = 0;
    // sync-free chunk sweeps (one-part meshes with wide levels)
    fh::DBuf<unsigned int> d_rowFlags;     // per chunk
    fh::DBuf<unsigned long long> d_rowTicket;
    fh::DBuf<unsigned int> d_rowErr;
    unsigned long long rowTicketBase = 0, rowEpoch = 0, rowWaitTicks = 0;
    bool rowSweep = false;
    // output rows of the value-polling chunk passes that run in place
    fh::DBuf<double> d_chunkTmp;
    fh::DBuf<float> d_chunkTmp32;
    // tiled sweeps (hex boxes)
    fh::DBuf<unsigned int> d_tileFlags;
    fh::DBuf<unsigned long long> d_tileTicket;
    unsigned long long tileTicketBase = 0, tileEpoch = 0;

    // live timing of chosen kernels inside the solver loops: a pair of HIP
    // events around every launch of a kernel whose FH_KERNEL_* bit is set in
    // mask; the pairs of the iterations that ran are summed after each batch
    struct KernelTiming
    {
        uint32_t mask = 0;
        std::vector<hipEvent_t> ev[5];
        int used[5] = {0, 0, 0, 0, 0};
        double ms[5] = {0, 0, 0, 0, 0};
        int64_t launches[5] = {0, 0, 0, 0, 0};
    } kt;

    // device scalar state and reduction partials
    fh::DBuf<fh::DevState> d_state;
    fh::DevState* h_state = nullptr;      // pinned
    fh::DBuf<double> d_partials;
    // arrival counter of the kernels that finalize themselves (zero
    // between launches)
    fh::DBuf<unsigned int> d_finCount;

    // work vectors (library numbering)
    fh::DBuf<double> v[10];
    fh::DBuf<float> f[10];
    fh::DBuf<double> sw[2];               // smoother residual / correction
    fh::DBuf<float> sw32[2];
    fh::DBuf<double> d_psiNat, d_srcNat;  // staging for host-array solves

    fh::RowsView rows() const
    {
        fh::RowsView R;
        R.n = mesh->nCells;
        R.rowInfo = mesh->d_rowInfo.p;
        R.sliceStart = mesh->d_sliceStart.p;
        R.col = mesh->d_col.p;
        R.face = mesh->d_face.p;
        if (mesh->nIfFaces > 0)
        {
            R.ifStart = mesh->d_ifStart.p;
            R.ifNbr = mesh->d_ifNbr.p;
            R.ifBase = mesh->nnzPad;
        }
        return R;
    }
    // length of the value arrays: the SELL entries, then the interface
    // entries of a coupled mesh
    int64_t nVal() const { return std::max<int64_t>(mesh->nnzPad + mesh->nIfFaces, 1); }

    fh::SchedView sched() const
    {
        const fh::Schedule& S = mesh->sched;
        fh::SchedView v;
        v.nParts = S.nParts;
        v.nLevels = S.nLevels;
        v.partLevelOff = S.d_partLevelOff.p;
        v.partMinLevel = S.d_partMinLevel.p;
        v.segStart = S.d_segStart.p;
        v.upOff = S.d_upOff.p;
        v.upList = S.d_upList.p;
        v.downOff = S.d_downOff.p;
        v.downList = S.d_downList.p;
        v.order = S.d_order.p;
        v.maxDeps = 0;
        return v;
    }

    fh::TileView tile_view(int bz, int bx = 16) const
    {
        const fh_mesh::TileShape* T = nullptr;
        for (const auto& t : mesh->tiles) if (t.bz == bz && t.bx == bx) T = &t;
        if (!T) fh::fail(FH_ERR_INVALID, "no tile shape %d x %d x %d", bx, bx, bz);
        fh::TileView v;
        v.nx = mesh->boxNx; v.ny = mesh->boxNy; v.nz = mesh->boxNz;
        v.bx = bx; v.by = bx; v.bz = bz;
        v.nTI = T->nTI; v.nTJ = T->nTJ; v.nTK = T->nTK;
        v.order = T->d_order.p;
        return v;
    }

    fh::TileSync tile_sync()
    {
        fh::TileSync s;
        s.flags = d_tileFlags.p;
        s.ticket = d_tileTicket.p;
        s.ticketBase = tileTicketBase;
        s.epoch = tileEpoch;
        return s;
    }

    fh::RowSync row_sync()
    {
        fh::RowSync s;
        s.flags = d_rowFlags.p;
        s.ticket = d_rowTicket.p;
        s.ticketBase = rowTicketBase;
        s.epoch = rowEpoch;
        s.waitTicks = rowWaitTicks;
        s.err = d_rowErr.p;
        return s;
    }

    fh::ChunkView chunks() const
    {
        fh::ChunkView c;
        c.nChunks = mesh->nChunks;
        c.start = mesh->d_chunkStart.p;
        c.fwdOff = mesh->d_fwdOff.p;
        c.fwdDeps = mesh->d_fwdDeps.p;
        c.bwdOff = mesh->d_bwdOff.p;
        c.bwdDeps = mesh->d_bwdDeps.p;
        return c;
    }

    fh::SweepSync sync()
    {
        fh::SweepSync s;
        s.ticket = d_sync.p;
        s.progress = d_sync.p + 1;
        s.epoch = epoch;
        return s;
    }
};

namespace fh
{
void matrix_init(fh_matrix* M, fh_mesh* mesh, hipStream_t shared = nullptr);
void matrix_free(fh_matrix* M);
void matrix_set(fh_matrix* M, const double* diag, const double* upper, const double* lower,
                bool device);
void matrix_coeffs_updated(fh_matrix* M);
void ensure_sumA(fh_matrix* M);
void ensure_valT(fh_matrix* M);
void ensure_f32(fh_matrix* M);
// factor of the DIC/FDIC (0), DILU (1), diagonal (2), tile-local DIC (3)
// or tile-local DILU (4) kind, FP64 or FP32
const double* ensure_rD(fh_matrix* M, int which);
const float* ensure_rD32(fh_matrix* M, int which);

// preconditioner apply w = M^-1 r on library-numbered vectors.  With
// partials != null it also leaves sum w*r in the partials and returns their
// block count (for launch_finalize_blocks).  Paths whose last kernel can
// finalize itself apply F there and set *fused.
template<class T>
int32_t apply_precond(fh_matrix* M, int precon, const T* r, T* w, double* partials,
                      DevState* st, bool transpose, const fh::FusedFinalize& F = {},
                      bool* fused = nullptr, double* tileSum = nullptr);

void profile_kernel(fh_matrix* M, int kernel, int precision, int precon, int reps,
                    double* ms, double* bytes);
void solve_device(fh_matrix* M, const fh_controls* c, double* d_psiNat,
                  const double* d_srcNat, fh_perf* perf);
void smooth_device(fh_matrix* M, int smoother, double* psi, const double* src,
                   int32_t nSweeps, DevState* st);
template<class T>
void smooth_t(fh_matrix* M, int smoother, T* psi, const T* src, int32_t nSweeps,
              DevState* st);
// pieces shared with the GAMG driver
void state_upload(fh_matrix* M, const fh_controls* c, double tol, double relTol,
                  int32_t maxIter, int32_t minIter);
void init_fp64(fh_matrix* M, const double* psi, const double* src, double* rA, double* tmp);
template<class T> const T* diag_for(fh_matrix* M);
template<class T> const T* val_for(fh_matrix* M, bool transpose);
template<> const double* diag_for<double>(fh_matrix* M);
template<> const float* diag_for<float>(fh_matrix* M);
template<> const double* val_for<double>(fh_matrix* M, bool transpose);
template<> const float* val_for<float>(fh_matrix* M, bool transpose);
template<class T>
void pcg_loop(fh_matrix* M, int precon, T* psi, T* r, T* p, T* w, int32_t maxIter);
const char* precon_name(int p);
void fill_perf(fh_matrix* M, fh_perf* perf, const char* name);
// FH_ERR_SCHEDULE when a chunk-sweep wait expired since the last check
void check_sweep_errors(fh_matrix* M);

// batched enqueue of device iterations with a poll of DevState.done: the
// batch doubles up to 32 iterations between host polls
// event around a launch of kernel id k (FH_KERNEL_*) when it is timed
inline void kt_mark(fh_matrix* M, int k)
{
    auto& T = M->kt;
    if (!(T.mask & (1u << k))) return;
    if (T.used[k] == (int)T.ev[k].size())
    {
        hipEvent_t e;
        FH_HIP(hipEventCreate(&e));
        T.ev[k].push_back(e);
    }
    FH_HIP(hipEventRecord(T.ev[k][T.used[k]++], M->stream));
}

// after a synchronised batch: add the pairs of the iterations that ran
// (launches after convergence exit at once and are not counted)
inline void kt_harvest(fh_matrix* M, int32_t itBefore, int32_t itAfter)
{
    auto& T = M->kt;
    for (int k = 0; k < 5; k++)
    {
        const int pairs = std::min(T.used[k]/2, std::max(0, itAfter - itBefore));
        for (int i = 0; i < pairs; i++)
        {
            float ms = 0.f;
            FH_HIP(hipEventElapsedTime(&ms, T.ev[k][2*i], T.ev[k][2*i + 1]));
            T.ms[k] += ms;
            T.launches[k]++;
        }
        T.used[k] = 0;
    }
}

template<class F>
void run_batches(fh_matrix* M, int32_t maxIter, F&& enqueue)
{
    int32_t batch = 1, seq = 0;
    const int64_t cap = 4*(int64_t)(maxIter > 1 ? maxIter : 1) + 64;
    for (;;)
    {
        const int32_t itBefore = M->h_state->nIterations;
        for (int32_t i = 0; i < batch; i++) enqueue(seq++);
        FH_HIP(hipMemcpyAsync(M->h_state, M->d_state.p, sizeof(DevState),
                              hipMemcpyDeviceToHost, M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
        if (M->kt.mask) kt_harvest(M, itBefore, M->h_state->nIterations);
        if (M->h_state->done) break;
        if (seq > cap) fail(FH_ERR_SCHEDULE, "solver did not terminate after %d steps", seq);
        batch = batch*2 < 32 ? batch*2 : 32;
    }
}
} // namespace fh
