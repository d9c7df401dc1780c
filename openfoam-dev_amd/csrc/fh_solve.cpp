// Matrix handling and the device solver drivers: PCG (PCG.C:65-193),
// PBiCGStab (PBiCGStab.C:68-254), PBiCG (PBiCG.C:65-222), smoothSolver
// (smoothSolver.C:77-193) and diagonalSolver, plus the mixed-precision
// iterative refinement around FP32 inner PCG / PBiCGStab solves.
#include "fh_matrix.h"
#include "fh_hops.h"
#include "fh_gamg.h"
#include "fh_gamg_kernels.h"
#include "fh_coarse.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

namespace fh
{

void matrix_init(fh_matrix* M, fh_mesh* mesh, hipStream_t shared)
{
    M->mesh = mesh;
    FH_HIP(hipSetDevice(mesh->device));
    if (shared)
    {
        M->stream = shared;
        M->ownsStream = false;
    }
    else
    {
        FH_HIP(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
        M->ownsStream = true;
    }
    FH_HIP(hipEventCreate(&M->evStart));
    FH_HIP(hipEventCreate(&M->evStop));
    M->d_state.alloc(1);
    FH_HIP(hipHostMalloc((void**)&M->h_state, sizeof(DevState), hipHostMallocDefault));
    std::memset(M->h_state, 0, sizeof(DevState));
    size_t nPart = 2*reduction_partials(mesh->nCells) + 64;
    if (mesh->tiled)
    {
        nPart = std::max(nPart, 2*(size_t)box_spmv_blocks(mesh->boxNx, mesh->boxNy,
                                                          mesh->boxNz) + 64);
        // one partial per tile of the fused sweep sums
        nPart = std::max(nPart, 2*(size_t)mesh->maxTiles + 64);
        // the tile-local apply (one partial per 16^3 tile) followed by the
        // coarse correction's partials
        const size_t nT16 = (size_t)((mesh->boxNx + 15)/16)*((mesh->boxNy + 15)/16)
                            *((mesh->boxNz + 15)/16);
        nPart = std::max(nPart, 4*nT16 + 64);
    }
    // (nPart counts a and b partials; room for the group sums of the
    // self-finalizing kernels)
    M->d_partials.alloc(fin_partials(nPart/2));
    M->d_finCount.alloc(fin_counters(nPart/2));
    FH_HIP(hipMemsetAsync(M->d_finCount.p, 0, M->d_finCount.bytes(), M->stream));
    M->d_sync.alloc(1 + (size_t)mesh->sched.nParts);
    FH_HIP(hipMemsetAsync(M->d_sync.p, 0, M->d_sync.bytes(), M->stream));
    // one-part meshes sweep row by row unless the level sweep is forced
    const char* sw = std::getenv("FOAMHIP_SWEEP");
    M->rowSweep = mesh->chunkSweep && !(sw && std::strcmp(sw, "parts") == 0);
    if (M->rowSweep)
    {
        M->d_rowFlags.alloc(std::max(mesh->nChunks, 1));
        M->d_rowTicket.alloc(1);
        M->d_rowErr.alloc(1);
        FH_HIP(hipMemsetAsync(M->d_rowFlags.p, 0, M->d_rowFlags.bytes(), M->stream));
        FH_HIP(hipMemsetAsync(M->d_rowTicket.p, 0, M->d_rowTicket.bytes(), M->stream));
        FH_HIP(hipMemsetAsync(M->d_rowErr.p, 0, M->d_rowErr.bytes(), M->stream));
        // a legitimate wait lasts at most one sweep; 4 s of the device's
        // wall clock (kHz rate) marks a broken schedule
        int khz = 0;
        FH_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, mesh->device));
        M->rowWaitTicks = 4000ull*(unsigned long long)(khz > 0 ? khz : 100000);
    }
    if (mesh->tiled)
    {
        // FOAMHIP_TILE_BZ: tile depth of the FP64 global preconditioner
        // sweeps (4 or 8); FP32 ones run at the deeper default
        M->preBz64 = TileBZ<double>::pre;
        M->preBz32 = TileBZ<float>::pre;
        if (const char* e = std::getenv("FOAMHIP_TILE_BZ"))
        {
            const int v = std::atoi(e);
            if (v == TileBZ<double>::pre || v == TileBZ<double>::alt) M->preBz64 = v;
        }
        M->d_tileFlags.alloc(std::max(mesh->maxTiles, 1));
        M->d_tileTicket.alloc(1);
        FH_HIP(hipMemsetAsync(M->d_tileFlags.p, 0, M->d_tileFlags.bytes(), M->stream));
        FH_HIP(hipMemsetAsync(M->d_tileTicket.p, 0, M->d_tileTicket.bytes(), M->stream));
    }
    FH_HIP(hipStreamSynchronize(M->stream));
    M->epoch = 0;
}

void matrix_free(fh_matrix* M)
{
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    gamg_release(M);
    if (M->cblas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(M->cblas));
    M->cblas = nullptr;
    if (M->h_state) (void)hipHostFree(M->h_state);
    M->h_state = nullptr;
    if (M->evStart) (void)hipEventDestroy(M->evStart);
    if (M->evStop) (void)hipEventDestroy(M->evStop);
    for (auto& v : M->kt.ev)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    if (M->stream && M->ownsStream) (void)hipStreamDestroy(M->stream);
    M->stream = nullptr;
}

void matrix_set(fh_matrix* M, const double* diag, const double* upper, const double* lower,
                bool device)
{
    fh_mesh* mesh = M->mesh;
    const int32_t N = mesh->nCells, F = mesh->nFaces;
    if (!diag) fail(FH_ERR_INVALID, "diag must not be NULL");
    if (!upper && lower) fail(FH_ERR_INVALID, "lower given without upper");
    hipStream_t s = M->stream;
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;

    M->diagonalOnly = (upper == nullptr);
    M->symmetric = (lower == nullptr);
    DBuf<double> diagNat;
    diagNat.alloc(N);
    FH_HIP(hipMemcpyAsync(diagNat.p, diag, sizeof(double)*N, kind, s));
    M->d_diag.alloc(N);
    launch_permute_in<double, double>(N, mesh->d_perm.p, diagNat.p, M->d_diag.p, s);
    if (!M->diagonalOnly && F > 0)
    {
        M->d_upperF.alloc(F);
        FH_HIP(hipMemcpyAsync(M->d_upperF.p, upper, sizeof(double)*F, kind, s));
        if (!M->symmetric)
        {
            M->d_lowerF.alloc(F);
            FH_HIP(hipMemcpyAsync(M->d_lowerF.p, lower, sizeof(double)*F, kind, s));
        }
        else
        {
            M->d_lowerF.release();
        }
    }
    else
    {
        M->d_upperF.release();
        M->d_lowerF.release();
    }
    matrix_coeffs_updated(M);
}

// after d_upperF/d_lowerF/d_diag changed: rebuild the SELL values and drop
// every cached factor
void matrix_coeffs_updated(fh_matrix* M)
{
    fh_mesh* mesh = M->mesh;
    hipStream_t s = M->stream;
    M->d_val.alloc(M->nVal());
    const double* lo = M->symmetric ? M->d_upperF.p : M->d_lowerF.p;
    if (M->diagonalOnly || mesh->nFaces == 0)
        FH_HIP(hipMemsetAsync(M->d_val.p, 0, M->d_val.bytes(), s));
    else
        launch_gather_values<double>(M->rows(), mesh->nnzPad, M->d_upperF.p, lo, M->d_val.p, s);
    // the interface entries: -interfaceBouCoeffs (zero until they are set)
    if (mesh->nIfFaces > 0)
    {
        if (M->ifSet)
            launch_interface_values(mesh->nIfFaces, mesh->d_ifFace.p, M->d_ifBou.p,
                                    M->d_val.p + mesh->nnzPad, s);
        else
            FH_HIP(hipMemsetAsync(M->d_val.p + mesh->nnzPad, 0,
                                  sizeof(double)*mesh->nIfFaces, s));
    }
    FH_HIP(hipStreamSynchronize(s));
    M->hasCoeffs = true;
    M->valTValid = M->sumAValid = M->f32Valid = M->dir32Valid = M->dir64Valid = false;
    M->dir16Valid = false;
    M->padValid = false;
    M->halfState = 0;
    M->coarseState = 0;
    for (int i = 0; i < 5; i++) M->rDValid[i] = M->rD32Valid[i] = false;
    gamg_invalidate(M);
}

void ensure_sumA(fh_matrix* M)
{
    if (M->sumAValid) return;
    M->d_sumA.alloc(M->mesh->nCells);
    launch_sum_a(M->rows(), M->d_diag.p, M->d_val.p, M->d_sumA.p, M->stream);
    M->sumAValid = true;
}

void ensure_valT(fh_matrix* M)
{
    if (M->valTValid) return;
    fh_mesh* mesh = M->mesh;
    M->d_valT.alloc(M->nVal());
    if (M->diagonalOnly || mesh->nFaces == 0)
        FH_HIP(hipMemsetAsync(M->d_valT.p, 0, M->d_valT.bytes(), M->stream));
    else
    {
        const double* lo = M->symmetric ? M->d_upperF.p : M->d_lowerF.p;
        // swapped roles: lower entries take upper[f], upper entries lower[f]
        launch_gather_values<double>(M->rows(), mesh->nnzPad, lo, M->d_upperF.p,
                                     M->d_valT.p, M->stream);
    }
    // Tmul's interfaces take the interfaceIntCoeffs (lduMatrixATmul.C:95-151)
    if (mesh->nIfFaces > 0)
    {
        if (M->ifSet)
            launch_interface_values(mesh->nIfFaces, mesh->d_ifFace.p, M->d_ifInt.p,
                                    M->d_valT.p + mesh->nnzPad, M->stream);
        else
            FH_HIP(hipMemsetAsync(M->d_valT.p + mesh->nnzPad, 0,
                                  sizeof(double)*mesh->nIfFaces, M->stream));
    }
    M->valTValid = true;
}

void ensure_f32(fh_matrix* M)
{
    if (M->f32Valid) return;
    fh_mesh* mesh = M->mesh;
    M->d_diag32.alloc(mesh->nCells);
    M->d_val32.alloc(M->nVal());
    launch_convert<double, float>(mesh->nCells, M->d_diag.p, M->d_diag32.p, M->stream);
    launch_convert<double, float>(M->nVal(), M->d_val.p, M->d_val32.p, M->stream);
    M->f32Valid = true;
}

// tile depth of the global preconditioner sweeps on a box
template<class T> static int pre_bz(const fh_matrix* M);
template<> int pre_bz<double>(const fh_matrix* M) { return M->preBz64; }
template<> int pre_bz<float>(const fh_matrix* M) { return M->preBz32; }

template<class T> static DBuf<T>& chunk_tmp(fh_matrix* M);
template<> DBuf<double>& chunk_tmp<double>(fh_matrix* M) { return M->d_chunkTmp; }
template<> DBuf<float>& chunk_tmp<float>(fh_matrix* M) { return M->d_chunkTmp32; }

template<class T>
static void sweep_any(fh_matrix* M, int kind, const SweepData<T>& d, DevState* st,
                      bool local = false)
{
    if (M->mesh->tiled)
    {
        TileSync ts = M->tile_sync();
        const bool pre = kind == SWEEP_PRECOND_FWD || kind == SWEEP_PRECOND_BWD
                      || kind == SWEEP_PRECOND_BWD_T;
        const int bz = pre && !local ? pre_bz<T>(M) : tile_bz<T>(kind);
        launch_tile_sweep<T>(kind, M->tile_view(bz), M->rows(), d, ts, st, M->stream, local);
        M->tileTicketBase = ts.ticketBase;
        M->tileEpoch = ts.epoch;
        return;
    }
    // one-part meshes (GAMG coarse levels, unstructured meshes): the
    // multi-workgroup chunk sweep where the mesh has chunks (levels of 16
    // rows or more on average), with each row's operands loaded before its
    // chunk's wait, so a dependency level costs one flag hand-off and one
    // round of value loads (GAMG levels of 7k-2M rows: 1.0-3.2 ms per
    // Gauss-Seidel sweep against 1.5-43 ms for one workgroup);
    // narrower dependency graphs (chains) are walked by one workgroup.
    // FOAMHIP_SWEEP=chunks or parts forces the multi-workgroup sweeps.
    static const bool multiWg = [] {
        const char* e = std::getenv("FOAMHIP_SWEEP");
        return e && (std::strcmp(e, "chunks") == 0 || std::strcmp(e, "parts") == 0);
    }();
    const fh_mesh* mesh = M->mesh;
    if (!multiWg && !mesh->chunkSweep && mesh->sched.nParts == 1 && mesh->nFaces > 0
        && launch_sweep_level_wg<T>(kind, M->rows(), M->sched(), d, st, M->stream))
        return;
    if (M->rowSweep)
    {
        RowSync rs = M->row_sync();
        // the Gauss-Seidel and DIC/DILU passes poll the dependency values
        // (three configuration-4 FP64 V-cycles: 0.146 s against 0.163 s with
        // chunk flags); FOAMHIP_CHUNK_SYNC=flags waits on the chunk flags
        static const bool values = [] {
            const char* e = std::getenv("FOAMHIP_CHUNK_SYNC");
            return !(e && std::strcmp(e, "flags") == 0);
        }();
        T* tmp = nullptr;
        if (values)
        {
            DBuf<T>& b = chunk_tmp<T>(M);
            b.alloc(mesh->nCells);
            tmp = b.p;
        }
        launch_sweep_rows<T>(kind, M->rows(), d, M->chunks(), rs, st, M->stream, mesh->nCells,
                             tmp);
        M->rowTicketBase = rs.ticketBase;
        M->rowEpoch = rs.epoch;
        return;
    }
    SweepSync sy = M->sync();
    launch_sweep<T>(kind, M->rows(), M->sched(), sy, d, st, M->stream);
    M->epoch = sy.epoch;
}

static void sweep_d(fh_matrix* M, int kind, const SweepData<double>& d, DevState* st,
                    bool local = false)
{
    sweep_any<double>(M, kind, d, st, local);
}

static void sweep_f(fh_matrix* M, int kind, const SweepData<float>& d, DevState* st,
                    bool local = false)
{
    sweep_any<float>(M, kind, d, st, local);
}

const double* ensure_rD(fh_matrix* M, int which)
{
    // tile-local factors exist on tiled boxes only; elsewhere the global one
    if (which >= 3 && !M->mesh->tiled) which -= 3;
    DBuf<double>& b = M->d_rD[which];
    if (M->rDValid[which]) return b.p;
    const int32_t N = M->mesh->nCells;
    b.alloc(N);
    if (which == 2 || M->diagonalOnly || M->mesh->nFaces == 0)
    {
        // diagonalPreconditioner.C:50-63 (and the factor of a matrix
        // without off-diagonals): 1/diag
        FH_HIP(hipMemcpyAsync(b.p, M->d_diag.p, sizeof(double)*N, hipMemcpyDeviceToDevice,
                              M->stream));
    }
    else
    {
        SweepData<double> d{};
        d.val = M->d_val.p;
        d.diag = M->d_diag.p;
        d.w = b.p;
        d.upperF = M->d_upperF.p;
        const bool dilu = which == 1 || which == 4;
        sweep_d(M, dilu ? SWEEP_RD_DILU : SWEEP_RD_DIC, d, nullptr, which >= 3);
    }
    launch_reciprocal<double>(N, b.p, M->stream);
    M->rDValid[which] = true;
    return b.p;
}

// The FP32 tile-local DIC of a symmetric box stores its off-diagonal
// coefficients in FP16 (rows that are not whole 16-wide vectors run on
// zero-padded rows): the factor and both
// passes use the coefficients rounded to FP16 precision (the incomplete
// Cholesky factor of that matrix; diagonal, factor and arithmetic FP32).
// They are stored scaled by 2^e, e chosen from the largest magnitude so
// that it lands just below 2^14; if the smallest non-zero magnitude would
// then fall below the FP16 normal range (2^-14, a span of 2^28), the
// coefficients stay FP32.  Decided once per coefficient update.
static constexpr int HALF_TOP = 14;      // largest scaled magnitude < 2^14
static constexpr int HALF_BOTTOM = -14;  // smallest FP16 normal 2^-14

static bool tile_dic_half(fh_matrix* M)
{
    if (!(M->symmetric && M->mesh->tiled)) return false;
    if (M->halfState == 0)
    {
        ensure_f32(M);
        DBuf<unsigned int> range;
        range.alloc(2);
        const unsigned int init[2] = {0u, 0xffffffffu};
        FH_HIP(hipMemcpyAsync(range.p, init, sizeof(init), hipMemcpyHostToDevice, M->stream));
        launch_abs_range(std::max<int64_t>(M->mesh->nnzPad, 1), M->d_val32.p, range.p,
                         M->stream);
        unsigned int got[2];
        FH_HIP(hipMemcpyAsync(got, range.p, sizeof(got), hipMemcpyDeviceToHost, M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
        float mx, mn;
        std::memcpy(&mx, &got[0], 4);
        std::memcpy(&mn, &got[1], 4);
        M->halfState = -1;
        if (got[0] == 0u)
        {
            // no off-diagonal coefficient: any scale
            M->halfState = 1;
            M->halfScale = 1.f;
        }
        else if (std::isfinite(mx))
        {
            int ex;
            std::frexp(mx, &ex);              // mx in [2^(ex-1), 2^ex)
            const int e = HALF_TOP - ex;      // mx*2^e < 2^14
            if (std::ldexp((double)mn, e) >= std::ldexp(1.0, HALF_BOTTOM) && std::abs(e) < 100)
            {
                M->halfState = 1;
                M->halfScale = (float)std::ldexp(1.0, e);
            }
        }
    }
    return M->halfState == 1;
}

const float* ensure_rD32(fh_matrix* M, int which)
{
    if (which >= 3 && !M->mesh->tiled) which -= 3;
    DBuf<float>& b = M->d_rD32[which];
    if (M->rD32Valid[which]) return b.p;
    const int32_t N = M->mesh->nCells;
    b.alloc(N);
    if (which >= 3 && !M->diagonalOnly && M->mesh->nFaces > 0)
    {
        // tile-local factor over the tiles of the FP32 sweeps (deeper than
        // the FP64 ones), so that factor and sweeps share their blocks
        ensure_f32(M);
        SweepData<float> d{};
        d.val = M->d_val32.p;
        d.diag = M->d_diag32.p;
        d.w = b.p;
        d.upperF = M->d_upperF.p;
        DBuf<float> rounded;
        if (which == 3 && tile_dic_half(M))
        {
            const int64_t nnz = std::max<int64_t>(M->mesh->nnzPad, 1);
            rounded.alloc(nnz);
            launch_round_half(nnz, M->d_val32.p, rounded.p, M->halfScale, M->stream);
            d.val = rounded.p;
        }
        sweep_f(M, which == 4 ? SWEEP_RD_DILU : SWEEP_RD_DIC, d, nullptr, true);
        launch_reciprocal<float>(N, b.p, M->stream);
        if (rounded.p) FH_HIP(hipStreamSynchronize(M->stream));
    }
    else
    {
        const double* r64 = ensure_rD(M, which);
        launch_convert<double, float>(N, r64, b.p, M->stream);
    }
    M->rD32Valid[which] = true;
    return b.p;
}

static int factor_kind(int precon)
{
    switch (precon)
    {
        case FH_PRECON_DIC:
        case FH_PRECON_FDIC: return 0;
        case FH_PRECON_DILU: return 1;
        case FH_PRECON_DIAGONAL: return 2;
        case FH_PRECON_TILE_DIC: return 3;
        case FH_PRECON_TILE_DILU: return 4;
        default: return -1;
    }
}

static void ensure_dir32(fh_matrix* M)
{
    if (M->dir32Valid) return;
    ensure_f32(M);
    const int32_t N = M->mesh->nCells;
    float* dst[6];
    for (int d = 0; d < 6; d++)
    {
        M->d_dir32[d].alloc(N);
        dst[d] = M->d_dir32[d].p;
    }
    launch_pack_dir_coefs<float>(M->tile_view(TileBZ<float>::pre), M->rows(), M->d_val32.p, 6,
                                 dst, M->stream);
    M->dir32Valid = true;
}

// FP64: the lower three of a symmetric matrix, all six otherwise
static void ensure_dir64(fh_matrix* M)
{
    if (M->dir64Valid) return;
    const int32_t N = M->mesh->nCells;
    const int nDir = M->symmetric ? 3 : 6;
    double* dst[6] = {};
    for (int d = 0; d < nDir; d++)
    {
        M->d_dir64[d].alloc(N);
        dst[d] = M->d_dir64[d].p;
    }
    for (int d = nDir; d < 6; d++) M->d_dir64[d].release();
    launch_pack_dir_coefs<double>(M->tile_view(TileBZ<double>::pre), M->rows(), M->d_val.p, nDir,
                                  dst, M->stream);
    M->dir64Valid = true;
}

// the pass's packed coefficients (k, j, i) for the global box sweeps
template<class T>
static void set_sweep_dirs(fh_matrix* M, SweepData<T>& d, bool backward);
template<>
void set_sweep_dirs<double>(fh_matrix* M, SweepData<double>& d, bool backward)
{
    ensure_dir64(M);
    if (!backward)
    {
        for (int q = 0; q < 3; q++) d.dir[q] = M->d_dir64[q].p;
        d.dirShift = 0;
    }
    else if (M->symmetric)
    {
        for (int q = 0; q < 3; q++) d.dir[q] = M->d_dir64[q].p;
        d.dirShift = 1;
    }
    else
    {
        d.dir[0] = M->d_dir64[5].p; d.dir[1] = M->d_dir64[4].p; d.dir[2] = M->d_dir64[3].p;
        d.dirShift = 0;
    }
}
template<>
void set_sweep_dirs<float>(fh_matrix* M, SweepData<float>& d, bool backward)
{
    ensure_dir32(M);
    if (!backward)
    {
        for (int q = 0; q < 3; q++) d.dir[q] = M->d_dir32[q].p;
    }
    else
    {
        d.dir[0] = M->d_dir32[5].p; d.dir[1] = M->d_dir32[4].p; d.dir[2] = M->d_dir32[3].p;
    }
    d.dirShift = 0;
}

template<class T> static const T* rD_for(fh_matrix* M, int which);
template<> const double* rD_for<double>(fh_matrix* M, int which) { return ensure_rD(M, which); }
template<> const float* rD_for<float>(fh_matrix* M, int which) { return ensure_rD32(M, which); }

template<> const double* val_for<double>(fh_matrix* M, bool transpose)
{
    if (transpose) { ensure_valT(M); return M->d_valT.p; }
    return M->d_val.p;
}
template<> const float* val_for<float>(fh_matrix* M, bool)
{
    ensure_f32(M);
    return M->d_val32.p;
}

template<> const double* diag_for<double>(fh_matrix* M) { return M->d_diag.p; }
template<> const float* diag_for<float>(fh_matrix* M) { ensure_f32(M); return M->d_diag32.p; }

template<class T>
static void sweep_t(fh_matrix* M, int kind, const SweepData<T>& d, DevState* st,
                    bool local = false);
template<> void sweep_t<double>(fh_matrix* M, int kind, const SweepData<double>& d,
                                DevState* st, bool local)
{
    sweep_d(M, kind, d, st, local);
}
template<> void sweep_t<float>(fh_matrix* M, int kind, const SweepData<float>& d,
                               DevState* st, bool local)
{
    sweep_f(M, kind, d, st, local);
}

// One global DIC / DILU pass on a box from packed coefficients (partials:
// the backward pass also sums w*x into one partial per tile); returns the
// partial count.  One-wavefront 8 x 8 x 16 tiles unless the workgroup
// kernel over 16 x 16 x bz tiles is asked for (FOAMHIP_BOX_SWEEP=wide).
template<class T>
static int32_t box_precond_pass(fh_matrix* M, bool backward, const SweepData<T>& d0,
                                double* partials, DevState* st,
                                const FusedFinalize& F = FusedFinalize(), bool* fused = nullptr)
{
    SweepData<T> d = d0;
    set_sweep_dirs<T>(M, d, backward);
    if (backward && partials)
    {
        d.dotR = d.x;
        d.partials = partials;
    }
    static const bool wide = [] {
        const char* e = std::getenv("FOAMHIP_BOX_SWEEP");
        return e && std::strcmp(e, "wide") == 0;
    }();
    if (!wide)
    {
        TileSync ts = M->tile_sync();
        const bool fuse = backward && partials && F.counter;
        const int32_t n = launch_sweep_wave<T>(backward, M->tile_view(16, 8), d, ts, st,
                                               M->stream, fuse ? F : FusedFinalize());
        if (fused) *fused = fuse;
        M->tileTicketBase = ts.ticketBase;
        M->tileEpoch = ts.epoch;
        return n;
    }
    sweep_t<T>(M, backward ? SWEEP_PRECOND_BWD : SWEEP_PRECOND_FWD, d, st, false);
    const TileView G = M->tile_view(pre_bz<T>(M));
    return G.nTI*G.nTJ*G.nTK;
}

template<class T>
int32_t apply_precond(fh_matrix* M, int precon, const T* r, T* w, double* partials,
                      DevState* st, bool transpose, const FusedFinalize& F, bool* fused,
                      double* tileSum)
{
    const int32_t N = M->mesh->nCells;
    if (fused) *fused = false;
    if (M->diagonalOnly || M->mesh->nFaces == 0)
    {
        // lduMatrixPreconditioner.C:84-93: a diagonal matrix gets no
        // preconditioner; DIC/DILU factors reduce to 1/diag otherwise
        if (precon == FH_PRECON_DIC || precon == FH_PRECON_FDIC || precon == FH_PRECON_DILU
            || precon == FH_PRECON_TILE_DIC || precon == FH_PRECON_TILE_DILU)
            precon = FH_PRECON_DIAGONAL;
        if (M->diagonalOnly) precon = FH_PRECON_NONE;
    }
    if (precon == FH_PRECON_GAMG)
    {
        // GAMGPreconditioner.C:87-148; PBiCG's transposed apply uses the
        // same V-cycles (GAMGPreconditioner.H:125-133)
        gamg_apply<T>(M, r, w, st);
        if (partials) launch_dot<T, T>(N, w, r, partials, st, M->stream);
        return reduction_blocks(N);
    }
    if (precon == FH_PRECON_NONE || precon == FH_PRECON_DIAGONAL)
    {
        const T* rD = precon == FH_PRECON_DIAGONAL ? rD_for<T>(M, 2) : nullptr;
        launch_precond_point<T>(N, precon, rD, r, w, partials ? partials : M->d_partials.p,
                                st, M->stream);
        return reduction_blocks(N);
    }
    const int which = factor_kind(precon);
    if (which < 0) fail(FH_ERR_UNSUPPORTED, "preconditioner %d has no sweep form", precon);
    SweepData<T> d{};
    d.val = val_for<T>(M, transpose);
    d.rD = rD_for<T>(M, which);
    d.x = r;
    d.w = w;
    d.perm = M->mesh->d_perm.p;
    const bool local = which >= 3 && M->mesh->tiled;
    const bool dilu = which == 1 || which == 4;
    if constexpr (std::is_same<T, float>::value)
    {
        if (local)
        {
            // bandwidth path of the FP32 tile-local factor: both passes in
            // one launch where the rows allow it
            ensure_dir32(M);
            const TileView G = M->tile_view(TileBZ<float>::pre);
            {
                const float* dir[6];
                for (int q = 0; q < 6; q++) dir[q] = M->d_dir32[q].p;
                const __half* h3[3] = {};
                const bool half = which == 3 && tile_dic_half(M);
                if (half && !M->dir16Valid && (G.nx & 15) == 0)
                {
                    __half* dst[3];
                    for (int q = 0; q < 3; q++) { M->d_dir16[q].alloc(N); dst[q] = M->d_dir16[q].p; }
                    launch_pack_half(N, dir, dst, M->halfScale, M->stream);
                    M->dir16Valid = true;
                }
                for (int q = 0; q < 3; q++) h3[q] = M->d_dir16[q].p;
                const bool fuse = partials && F.counter;
                const int32_t nb = launch_tile_local_apply_f32(G, r, d.rD, dir, M->symmetric, w,
                                                               st, M->stream, partials,
                                                               half ? h3 : nullptr,
                                                               fuse ? F : FusedFinalize(),
                                                               tileSum, M->halfScale);
                if (nb >= 0)
                {
                    if (fused) *fused = fuse;
                    return nb;
                }
            }
            {
                // rows not a multiple of 16: the fused kernels on rows padded
                // with zero cells (zero coefficients and factor), x and w
                // staged through padded copies; the same tiles, partials and
                // tile sums, the pad cells contributing zero
                const int32_t nx = G.nx, nxp = (nx + 15) & ~15;
                const int64_t rows = (int64_t)G.ny*G.nz, np = (int64_t)nxp*rows;
                const bool half = which == 3 && tile_dic_half(M);
                if (!M->padValid)
                {
                    for (int q = 0; q < 6; q++)
                    {
                        M->d_padDir32[q].alloc(np);
                        launch_pad_rows(nx, nxp, rows, M->d_dir32[q].p, M->d_padDir32[q].p,
                                        M->stream);
                    }
                    M->d_padRD.alloc(np);
                    launch_pad_rows(nx, nxp, rows, d.rD, M->d_padRD.p, M->stream);
                    if (half)
                    {
                        const float* src[3];
                        __half* dst[3];
                        for (int q = 0; q < 3; q++)
                        {
                            M->d_padDir16[q].alloc(np);
                            src[q] = M->d_padDir32[q].p;
                            dst[q] = M->d_padDir16[q].p;
                        }
                        launch_pack_half(np, src, dst, M->halfScale, M->stream);
                    }
                    M->d_padX.alloc(np);
                    M->d_padW.alloc(np);
                    M->padValid = true;
                }
                TileView Gp = G;
                Gp.nx = nxp;
                const float* dir[6];
                for (int q = 0; q < 6; q++) dir[q] = M->d_padDir32[q].p;
                const __half* h3[3];
                for (int q = 0; q < 3; q++) h3[q] = M->d_padDir16[q].p;
                const bool fuse = partials && F.counter;
                if (!M->padXFresh) launch_pad_rows(nx, nxp, rows, r, M->d_padX.p, M->stream);
                const int32_t nb = launch_tile_local_apply_f32(Gp, M->d_padX.p, M->d_padRD.p,
                                                               dir, M->symmetric, M->d_padW.p,
                                                               st, M->stream, partials,
                                                               half ? h3 : nullptr,
                                                               fuse ? F : FusedFinalize(),
                                                               tileSum, M->halfScale);
                if (nb < 0) fail(FH_ERR_INVALID, "internal: padded tile-local shape");
                M->padLeft = M->padKeep;
                if (!M->padKeep) launch_unpad_rows(nx, nxp, rows, M->d_padW.p, w, M->stream);
                if (fused) *fused = fuse;
                return nb;
            }
        }
    }
    if (M->mesh->tiled && !local && !transpose)
    {
        // global sweeps on a box: packed coefficients, and the backward pass
        // sums w*r per tile
        box_precond_pass<T>(M, false, d, nullptr, st);
        return box_precond_pass<T>(M, true, d, partials, st, F, fused);
    }
    sweep_t<T>(M, SWEEP_PRECOND_FWD, d, st, local);
    sweep_t<T>(M, transpose && dilu ? SWEEP_PRECOND_BWD_T : SWEEP_PRECOND_BWD, d, st, local);
    if (partials) launch_dot<T, T>(N, w, r, partials, st, M->stream);
    return reduction_blocks(N);
}

template int32_t apply_precond<double>(fh_matrix*, int, const double*, double*, double*, DevState*,
                                       bool, const FusedFinalize&, bool*, double*);
template int32_t apply_precond<float>(fh_matrix*, int, const float*, float*, double*, DevState*,
                                      bool, const FusedFinalize&, bool*, double*);

void state_upload(fh_matrix* M, const fh_controls* c, double tol, double relTol,
                         int32_t maxIter, int32_t minIter)
{
    DevState* h = M->h_state;
    std::memset(h, 0, sizeof(DevState));
    h->tolerance = tol;
    h->relTol = relTol;
    h->maxIter = maxIter;
    h->minIter = minIter;
    h->wArA = 1e20;      // SolverPerformance::great_ (PCG.C:92)
    h->earlySeq = -1;
    FH_HIP(hipMemcpyAsync(M->d_state.p, h, sizeof(DevState), hipMemcpyHostToDevice, M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
    (void)c;
}

static void state_upload_outer(fh_matrix* M, const DevState& outer)
{
    *M->h_state = outer;
    M->h_state->done = 0;
    FH_HIP(hipMemcpyAsync(M->d_state.p, M->h_state, sizeof(DevState), hipMemcpyHostToDevice,
                          M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
}

// FP64 initialisation shared by the solvers: wA = A psi, normFactor,
// rA = source - wA, initial residual (lduMatrixSolver.C:174-197)
void init_fp64(fh_matrix* M, const double* psi, const double* src, double* rA,
                      double* tmp)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    ensure_sumA(M);
    launch_amul<double>(M->rows(), M->d_diag.p, M->d_val.p, psi, tmp, nullptr, M->stream);
    launch_sum_abs<double>(N, psi, part, M->stream);
    launch_finalize(N, part, FIN_XREF, st, 0, M->stream);
    launch_init_residual(N, tmp, src, M->d_sumA.p, rA, part, st, M->stream);
    launch_finalize(N, part, FIN_INIT, st, 0, M->stream);
}

// the finalize of op applied by the last block of the reduction kernel
// itself (FOAMHIP_FUSED_FINALIZE=0: separate finalize launches)
static FusedFinalize fused_fin(fh_matrix* M, int32_t op)
{
    static const bool off = [] {
        const char* e = std::getenv("FOAMHIP_FUSED_FINALIZE");
        return e && std::strcmp(e, "0") == 0;
    }();
    FusedFinalize F;
    if (off) return F;
    F.counter = M->d_finCount.p;
    F.st = M->d_state.p;
    F.op = op;
    F.n = M->mesh->nCells;
    return F;
}

// preconditioner apply with sum w*r, finalized by op (inside the apply's
// last kernel where its path allows)
template<class T>
static void precond_finalize(fh_matrix* M, int precon, const T* r, T* w, int32_t op)
{
    bool fused = false;
    const int32_t nb = apply_precond<T>(M, precon, r, w, M->d_partials.p, M->d_state.p, false,
                                        fused_fin(M, op), &fused);
    if (!fused)
        launch_finalize_blocks(nb, M->mesh->nCells, M->d_partials.p, op, M->d_state.p, 0,
                               M->stream);
}

template<class T>
void pcg_loop(fh_matrix* M, int precon, T* psi, T* r, T* p, T* w, int32_t maxIter)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    const T* diag = diag_for<T>(M);
    const T* val = val_for<T>(M, false);
    if (precon != FH_PRECON_NONE && factor_kind(precon) >= 0)
        (void)rD_for<T>(M, factor_kind(precon));
    if constexpr (std::is_same<T, double>::value)
    {
        const fh_mesh* mesh = M->mesh;
        if (mesh->tiled && (mesh->boxNx & 3) == 0)
        {
            // boxes: p = w + beta*p fused into the direction-packed product
            // (two p buffers in turn, A p in its own buffer); same values as
            // the separate kernels below
            ensure_dir64(M);
            const double* dir[6];
            for (int q = 0; q < 6; q++) dir[q] = M->d_dir64[q].p;
            M->v[7].alloc(N);
            M->v[8].alloc(N);
            double* pBuf[2] = {p, M->v[7].p};
            double* ap = M->v[8].p;
            run_batches(M, maxIter, [&](int32_t seq)
            {
                kt_mark(M, FH_KERNEL_PRECONDITION);
                precond_finalize<T>(M, precon, r, w, FIN_PCG_WARA);
                kt_mark(M, FH_KERNEL_PRECONDITION);
                double* pNew = pBuf[(seq + 1) & 1];
                const FusedFinalize fa = fused_fin(M, FIN_PCG_WAPA);
                kt_mark(M, FH_KERNEL_SPMV);
                const int32_t nb = launch_box_pcg_spmv<double>(mesh->boxNx, mesh->boxNy,
                                                               mesh->boxNz, diag, dir,
                                                               M->symmetric, w, pBuf[seq & 1],
                                                               pNew, ap, part, st, M->stream,
                                                               fa);
                kt_mark(M, FH_KERNEL_SPMV);
                if (!fa.counter)
                    launch_finalize_blocks(nb, N, part, FIN_PCG_WAPA, st, 0, M->stream);
                const FusedFinalize fr = fused_fin(M, FIN_PCG_RES);
                kt_mark(M, FH_KERNEL_AXPY_NORM);
                launch_pcg_update<T>(N, psi, r, pNew, ap, part, st, M->stream, fr);
                kt_mark(M, FH_KERNEL_AXPY_NORM);
                if (!fr.counter) launch_finalize(N, part, FIN_PCG_RES, st, 0, M->stream);
            });
            return;
        }
    }
    run_batches(M, maxIter, [&](int32_t)
    {
        const int32_t nb = apply_precond<T>(M, precon, r, w, part, st, false);
        launch_finalize_blocks(nb, N, part, FIN_PCG_WARA, st, 0, M->stream);
        launch_pcg_p<T>(N, w, p, st, M->stream);
        launch_spmv_dot<T>(M->rows(), diag, val, p, w, p, nullptr, part, st, M->stream);
        launch_finalize(N, part, FIN_PCG_WAPA, st, 0, M->stream);
        launch_pcg_update<T>(N, psi, r, p, w, part, st, M->stream);
        launch_finalize(N, part, FIN_PCG_RES, st, 0, M->stream);
    });
}

template<class T>
static void bicgstab_loop(fh_matrix* M, int precon, T* psi, T* rA, T** V, int32_t maxIter)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    const T* diag = diag_for<T>(M);
    const T* val = val_for<T>(M, false);
    T *pA = V[0], *yA = V[1], *AyA = V[2], *sA = V[3], *zA = V[4], *tA = V[5], *rA0 = V[6];
    if (precon != FH_PRECON_NONE && factor_kind(precon) >= 0)
        (void)rD_for<T>(M, factor_kind(precon));
    FH_HIP(hipMemcpyAsync(rA0, rA, sizeof(T)*N, hipMemcpyDeviceToDevice, M->stream));
    launch_dot<T, T>(N, rA0, rA, part, st, M->stream);
    launch_finalize(N, part, FIN_BICG_START, st, 0, M->stream);
    run_batches(M, maxIter, [&](int32_t seq)
    {
        launch_bicg_p<T>(N, rA, pA, AyA, st, M->stream);
        apply_precond<T>(M, precon, pA, yA, nullptr, st, false);
        launch_spmv_dot<T>(M->rows(), diag, val, yA, AyA, rA0, nullptr, part, st, M->stream);
        launch_finalize(N, part, FIN_BICG_ALPHA, st, seq, M->stream);
        launch_bicg_s<T>(N, rA, AyA, sA, part, st, M->stream);
        launch_finalize(N, part, FIN_BICG_S, st, seq, M->stream);
        launch_bicg_early<T>(N, psi, yA, st, seq, M->stream);
        apply_precond<T>(M, precon, sA, zA, nullptr, st, false);
        launch_spmv_dot<T>(M->rows(), diag, val, zA, tA, nullptr, sA, part, st, M->stream);
        launch_finalize(N, part, FIN_BICG_OMEGA, st, seq, M->stream);
        launch_bicg_update<T>(N, psi, yA, zA, rA, sA, tA, rA0, part, st, M->stream);
        launch_finalize(N, part, FIN_BICG_RES, st, seq, M->stream);
    });
}

void fill_perf(fh_matrix* M, fh_perf* perf, const char* name)
{
    const DevState* h = M->h_state;
    std::memset(perf, 0, sizeof(*perf));
    std::snprintf(perf->solverName, sizeof(perf->solverName), "%s", name);
    perf->initialResidual = h->initialResidual;
    perf->finalResidual = h->finalResidual;
    perf->nIterations = h->nIterations;
    perf->converged = h->converged;
    perf->singular = h->singular;
    perf->normFactor = h->normFactor;
}

void check_sweep_errors(fh_matrix* M)
{
    if (!M->rowSweep) return;
    unsigned int e = 0;
    FH_HIP(hipMemcpyAsync(&e, M->d_rowErr.p, sizeof(e), hipMemcpyDeviceToHost, M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
    if (e)
    {
        FH_HIP(hipMemsetAsync(M->d_rowErr.p, 0, sizeof(e), M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
        fail(FH_ERR_SCHEDULE, "a sweep dependency wait expired");
    }
}

const char* precon_name(int p)
{
    switch (p)
    {
        case FH_PRECON_NONE: return "none";
        case FH_PRECON_DIAGONAL: return "diagonal";
        case FH_PRECON_DIC: return "DIC";
        case FH_PRECON_FDIC: return "FDIC";
        case FH_PRECON_DILU: return "DILU";
        case FH_PRECON_GAMG: return "GAMG";
        case FH_PRECON_TILE_DIC: return "tileDIC";
        case FH_PRECON_TILE_DILU: return "tileDILU";
        default: return "?";
    }
}

// validation against the reference run-time tables (lduMatrixSolver.C:62-120,
// lduMatrixPreconditioner.C:94-140)
static void validate_selection(const fh_matrix* M, const fh_controls* c)
{
    if (c->agglomerator != FH_AGGLOM_ALGEBRAIC_PAIR && c->agglomerator != FH_AGGLOM_FACE_AREA_PAIR)
        fail(FH_ERR_INVALID, "Unknown GAMGAgglomeration type %d; valid: algebraicPair faceAreaPair",
             c->agglomerator);
    if (M->diagonalOnly) return;
    const bool sym = M->symmetric;
    const int s = c->solver;
    if (sym && s == FH_SOLVER_PBICG)
        fail(FH_ERR_INVALID, "Unknown symmetric matrix solver PBiCG; valid: GAMG PBiCGStab PCG smoothSolver");
    if (!sym && s == FH_SOLVER_PCG)
        fail(FH_ERR_INVALID, "Unknown asymmetric matrix solver PCG; valid: GAMG PBiCG PBiCGStab smoothSolver");
    if (s == FH_SOLVER_PCG || s == FH_SOLVER_PBICGSTAB || s == FH_SOLVER_PBICG)
    {
        const int p = c->preconditioner;
        if (sym && p == FH_PRECON_DILU)
            fail(FH_ERR_INVALID, "Unknown symmetric matrix preconditioner DILU; valid: DIC FDIC GAMG diagonal none");
        if (!sym && (p == FH_PRECON_DIC || p == FH_PRECON_FDIC))
            fail(FH_ERR_INVALID, "Unknown asymmetric matrix preconditioner %s; valid: DILU GAMG diagonal none",
                 precon_name(p));
    }
    if (s == FH_SOLVER_SMOOTH || s == FH_SOLVER_GAMG)
    {
        const int m = c->smoother;
        if (sym && (m == FH_SMOOTH_DILU || m == FH_SMOOTH_DILUGS))
            fail(FH_ERR_INVALID, "Unknown symmetric matrix smoother; valid: DIC DICGaussSeidel FDIC GaussSeidel nonBlockingGaussSeidel symGaussSeidel");
        if (!sym && (m == FH_SMOOTH_DIC || m == FH_SMOOTH_DICGS || m == FH_SMOOTH_FDIC))
            fail(FH_ERR_INVALID, "Unknown asymmetric matrix smoother; valid: DILU DILUGaussSeidel GaussSeidel nonBlockingGaussSeidel symGaussSeidel");
    }
}

// ---------------------------------------------------------------------------
// smoothers (library numbering, FP64)
// ---------------------------------------------------------------------------
template<class T> static DBuf<T>* smooth_work(fh_matrix* M);
template<> DBuf<double>* smooth_work<double>(fh_matrix* M) { return M->sw; }
template<> DBuf<float>* smooth_work<float>(fh_matrix* M) { return M->sw32; }

// Greedy colouring in the caller's cell order: a cell takes the smallest
// colour none of its lower-numbered neighbours has (on a box in natural
// order: red-black); rows are grouped by colour, ascending within a colour.
static std::mutex g_colourMutex;
static void ensure_colours(fh_mesh* m)
{
    std::lock_guard<std::mutex> lk(g_colourMutex);
    if (m->coloured) return;
    const int32_t n = m->nCells;
    std::vector<int32_t> start(n + 1, 0), nb((size_t)m->nFaces);
    for (int32_t f = 0; f < m->nFaces; f++) start[m->upper[f] + 1]++;
    for (int32_t c = 0; c < n; c++) start[c + 1] += start[c];
    {
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int32_t f = 0; f < m->nFaces; f++) nb[fill[m->upper[f]]++] = m->lower[f];
    }
    std::vector<int32_t> colour(n, 0), seen(16, -1);
    int32_t nCol = 1;
    for (int32_t c = 0; c < n; c++)
    {
        for (int32_t q = start[c]; q < start[c + 1]; q++)
        {
            const int32_t k = colour[nb[q]];
            if (k >= (int32_t)seen.size()) seen.resize(k + 1, -1);
            seen[k] = c;
        }
        int32_t k = 0;
        while (k < (int32_t)seen.size() && seen[k] == c) k++;
        colour[c] = k;
        nCol = std::max(nCol, k + 1);
    }
    m->colourStart.assign(nCol + 1, 0);
    for (int32_t r = 0; r < n; r++) m->colourStart[colour[m->perm[r]] + 1]++;
    for (int32_t k = 0; k < nCol; k++) m->colourStart[k + 1] += m->colourStart[k];
    std::vector<int32_t> rows(n), fill(m->colourStart.begin(), m->colourStart.end() - 1);
    for (int32_t r = 0; r < n; r++) rows[fill[colour[m->perm[r]]]++] = r;
    m->maxColourRows = 0;
    for (int32_t k = 0; k < nCol; k++)
        m->maxColourRows = std::max(m->maxColourRows, m->colourStart[k + 1] - m->colourStart[k]);
    static const bool verbose = std::getenv("FOAMHIP_GAMG_VERBOSE") != nullptr;
    if (verbose)
        std::fprintf(stderr, "foamhip colours: %d cells, %d colours, largest %d rows\n", n, nCol,
                     m->maxColourRows);
    m->d_colourRows.upload(rows, m->setupStream);
    m->d_colourStart.upload(m->colourStart, m->setupStream);
    FH_HIP(hipStreamSynchronize(m->setupStream));
    m->coloured = true;
}

// GaussSeidelSmoother.C:66-187, symGaussSeidelSmoother.C:66-216,
// DICSmoother.C:67-116, DILUSmoother.C:67-119, DICGaussSeidelSmoother.C:79-89,
// DILUGaussSeidelSmoother.C:79-89 (library numbering)
// colours of at most this many rows (one per thread of a 1024-thread
// workgroup) are swept by one workgroup, consecutive ones in one launch
static constexpr int32_t kColourRunRows = 1024;

template<class T>
void smooth_t(fh_matrix* M, int smoother, T* psi, const T* src, int32_t nSweeps,
              DevState* st)
{
    const int32_t N = M->mesh->nCells;
    if (nSweeps <= 0 || smoother == FH_SMOOTH_NONE) return;
    const T* diag = diag_for<T>(M);
    const T* val = val_for<T>(M, false);
    auto gs = [&](bool sym)
    {
        SweepData<T> d{};
        d.val = val;
        d.diag = diag;
        d.x = src;
        d.w = psi;
        if constexpr (std::is_same<T, float>::value)
        {
            // FP32 (the V-cycles of the mixed GAMG): multi-colour
            // Gauss-Seidel, colours ascending (and descending for the
            // symmetric smoother); every colour is one parallel pass
            ensure_colours(M->mesh);
            const std::vector<int32_t>& cs = M->mesh->colourStart;
            const int32_t nCol = (int32_t)cs.size() - 1;
            const int32_t* rows = M->mesh->d_colourRows.p;
            // small meshes (at most one row per thread in every colour): all
            // sweeps in one workgroup launch
            if (M->mesh->maxColourRows <= kColourRunRows)
            {
                launch_sweep_colours_wg<T>(M->rows(), d, rows, M->mesh->d_colourStart.p, nCol,
                                           nSweeps, sym, st, M->stream);
                return;
            }
            // large colours one launch each; runs of consecutive colours of
            // at most kColourRunRows rows (the tail of the greedy colouring)
            // in one workgroup launch
            const int32_t* dcs = M->mesh->d_colourStart.p;
            auto small = [&](int32_t k) { return cs[k + 1] - cs[k] <= kColourRunRows; };
            auto pass = [&](bool desc)
            {
                int32_t q = 0;
                while (q < nCol)
                {
                    const int32_t k = desc ? nCol - 1 - q : q;
                    if (!small(k))
                    {
                        launch_sweep_colour<T>(M->rows(), d, rows + cs[k], cs[k + 1] - cs[k], st,
                                               M->stream);
                        q++;
                        continue;
                    }
                    int32_t e = q;
                    while (e < nCol && small(desc ? nCol - 1 - e : e)) e++;
                    // colours [q, e) in pass order
                    const int32_t k0 = desc ? nCol - e : q, k1 = desc ? nCol - q : e;
                    launch_sweep_colour_run<T>(M->rows(), d, rows, dcs, k0, k1, desc, st,
                                               M->stream);
                    q = e;
                }
            };
            for (int32_t s = 0; s < nSweeps; s++)
            {
                pass(false);
                if (sym) pass(true);
            }
            return;
        }
        if constexpr (std::is_same<T, double>::value)
        {
            if (M->mesh->nIfFaces > 0)
            {
                // coupled interfaces: each sweep starts from bPrime = source
                // less the interfaces' Jacobi term of psi before the sweep
                // (GaussSeidelSmoother.C:117-143; the symmetric smoother's
                // backward pass uses the same bPrime)
                DBuf<T>* wk = smooth_work<T>(M);
                wk[1].alloc(N);
                d.x = wk[1].p;
                for (int32_t s = 0; s < nSweeps; s++)
                {
                    launch_interface_bprime(M->rows(), val, psi, src, wk[1].p, M->stream);
                    sweep_t<T>(M, SWEEP_GS_FWD, d, st);
                    if (sym) sweep_t<T>(M, SWEEP_GS_BWD, d, st);
                }
                return;
            }
        }
        for (int32_t s = 0; s < nSweeps; s++)
        {
            sweep_t<T>(M, SWEEP_GS_FWD, d, st);
            if (sym) sweep_t<T>(M, SWEEP_GS_BWD, d, st);
        }
    };
    auto ilu = [&](bool dilu)
    {
        DBuf<T>* wk = smooth_work<T>(M);
        wk[0].alloc(N);
        wk[1].alloc(N);
        for (int32_t s = 0; s < nSweeps; s++)
        {
            launch_residual<T>(M->rows(), diag, val, psi, src, wk[0].p, st, M->stream);
            apply_precond<T>(M, dilu ? FH_PRECON_DILU : FH_PRECON_DIC, wk[0].p, wk[1].p,
                             nullptr, st, false);
            launch_axpy_gamg<T>(N, psi, wk[1].p, st, M->stream);
        }
    };
    switch (smoother)
    {
        case FH_SMOOTH_GS: gs(false); break;
        case FH_SMOOTH_SYMGS: gs(true); break;
        case FH_SMOOTH_DIC:
        case FH_SMOOTH_FDIC: ilu(false); break;
        case FH_SMOOTH_DILU: ilu(true); break;
        case FH_SMOOTH_DICGS: ilu(false); gs(false); break;
        case FH_SMOOTH_DILUGS: ilu(true); gs(false); break;
        default: fail(FH_ERR_INVALID, "unknown smoother %d", smoother);
    }
}

template void smooth_t<double>(fh_matrix*, int, double*, const double*, int32_t, DevState*);
template void smooth_t<float>(fh_matrix*, int, float*, const float*, int32_t, DevState*);

void smooth_device(fh_matrix* M, int smoother, double* psi, const double* src,
                   int32_t nSweeps, DevState* st)
{
    smooth_t<double>(M, smoother, psi, src, nSweeps, st);
}

// smoothSolver.C:77-193
static void smooth_solver(fh_matrix* M, const fh_controls* c, double* psi, const double* src)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    if (c->nSweeps < 0)
    {
        smooth_device(M, c->smoother, psi, src, -c->nSweeps, nullptr);
        FH_HIP(hipStreamSynchronize(M->stream));
        std::memset(M->h_state, 0, sizeof(DevState));
        M->h_state->nIterations = -c->nSweeps;
        return;
    }
    for (int i = 0; i < 2; i++) M->v[i].alloc(N);
    init_fp64(M, psi, src, M->v[0].p, M->v[1].p);
    run_batches(M, c->maxIter, [&](int32_t)
    {
        smooth_device(M, c->smoother, psi, src, c->nSweeps, st);
        launch_residual<double>(M->rows(), M->d_diag.p, M->d_val.p, psi, src, M->v[0].p, st,
                                M->stream);
        launch_sum_abs_state<double>(N, M->v[0].p, part, st, M->stream);
        launch_finalize_smooth(N, part, st, c->nSweeps, M->stream);
    });
}

// ---------------------------------------------------------------------------
// mixed precision PCG with reliable updates: the Krylov recurrence (DIC
// preconditioner, SpMV, search direction, recurrence residual) runs in
// FP32; the solution is accumulated in FP64, the dot products are FP64
// sums, and whenever the FP32 residual has fallen by a decade (or looks
// converged) the true residual b - A x is recomputed in FP64, tested with
// the reference criterion (normFactor-normalised L1) and replaces the FP32
// one.  The recurrence is never restarted, so the iteration count stays
// close to the FP64 one.
// ---------------------------------------------------------------------------
// The coarse correction of the tile-local inner preconditioner: the coarse
// matrix over the 16^3 tiles of a symmetric box (at most 4096 tiles, else
// the inner preconditioner stays one-level), factored (Cholesky of -Ac,
// rocSOLVER, FP64) and inverted once per coefficient update.
// FOAMHIP_TILE_COARSE=0 turns it off.
static bool ensure_coarse(fh_matrix* M)
{
    static const bool off = [] {
        const char* e = std::getenv("FOAMHIP_TILE_COARSE");
        return e && std::strcmp(e, "0") == 0;
    }();
    if (M->coarseState != 0) return M->coarseState > 0;
    M->coarseState = -1;
    const fh_mesh* mesh = M->mesh;
    if (off || !mesh->tiled || !M->symmetric) return false;
    const int32_t nTI = (mesh->boxNx + 15)/16, nTJ = (mesh->boxNy + 15)/16;
    const int32_t nTK = (mesh->boxNz + 15)/16;
    const int64_t nT = (int64_t)nTI*nTJ*nTK;
    if (nT < 2 || nT > 4096) return false;
    const int32_t n = (int32_t)nT, ld = (n + 3) & ~3;
    // room for the fine partials (one per tile) and the coarse kernel's
    // partials
    const int64_t nF = nT, nC = (nT + 3)/4;
    if (2*nF + (int64_t)fin_partials(nC) > (int64_t)M->d_partials.n
        || (int64_t)fin_counters(nC) > (int64_t)M->d_finCount.n)
        return false;
    hipStream_t s = M->stream;
    ensure_dir64(M);
    const double* dir[3] = {M->d_dir64[0].p, M->d_dir64[1].p, M->d_dir64[2].p};
    M->d_ctile.alloc(4*(size_t)n);
    M->d_cdense.alloc((size_t)n*n);
    M->d_cinv.alloc((size_t)n*ld);
    M->d_crc.alloc((size_t)n);
    M->d_cz.alloc((size_t)n);
    M->d_cinfo.alloc(2);
    launch_coarse_tiles(mesh->boxNx, mesh->boxNy, mesh->boxNz, nTI, nTJ, n, M->d_diag.p, dir,
                        M->d_ctile.p, s);
    // planes of at most 256 tiles: the block-tridiagonal inverse (Schur
    // complements over the tile planes, hand-written kernels); otherwise the
    // dense Cholesky of rocSOLVER.  A negative definite matrix (the
    // Laplacian) is inverted as -Ac: the kernels take the sign of the first
    // tile's diagonal.
    const int32_t P = nTI*nTJ;
    if (P <= 256 && nTK >= 2)
    {
        M->d_cU.alloc((size_t)nTK*P*P);
        M->d_cwork.alloc((size_t)coarse_block_work_doubles(n, P));
        FH_HIP(hipMemsetAsync(M->d_cinfo.p, 0, M->d_cinfo.bytes(), s));
        if (!launch_coarse_block_inverse(n, nTI, nTJ, ld, M->d_ctile.p, M->d_cU.p,
                                         M->d_cdense.p, M->d_cwork.p, M->d_cinv.p,
                                         M->d_cinfo.p, s))
            fail(FH_ERR_INVALID, "internal: coarse plane shape");
        int32_t bad = 0;
        FH_HIP(hipMemcpyAsync(&bad, M->d_cinfo.p, sizeof(bad), hipMemcpyDeviceToHost, s));
        FH_HIP(hipStreamSynchronize(s));
        if (bad) return false;                   // not definite: one-level
        M->cN = n;
        M->cLd = ld;
        M->coarseState = 1;
        return true;
    }
    if (!M->cblas)
    {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success)
            fail(FH_ERR_HIP, "rocblas_create_handle failed");
        M->cblas = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(M->cblas);
    if (rocblas_set_stream(h, s) != rocblas_status_success)
        fail(FH_ERR_HIP, "rocblas_set_stream failed");
    launch_coarse_dense(n, nTI, nTJ, M->d_ctile.p, M->d_cdense.p, s);
    if (rocsolver_dpotrf(h, rocblas_fill_lower, n, M->d_cdense.p, n, M->d_cinfo.p)
            != rocblas_status_success
        || rocsolver_dpotri(h, rocblas_fill_lower, n, M->d_cdense.p, n, M->d_cinfo.p + 1)
            != rocblas_status_success)
        fail(FH_ERR_HIP, "rocSOLVER Cholesky of the coarse matrix failed to launch");
    int32_t info[2] = {0, 0};
    FH_HIP(hipMemcpyAsync(info, M->d_cinfo.p, sizeof(info), hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    if (info[0] != 0 || info[1] != 0) return false;      // not definite: one-level
    launch_coarse_inv32(n, ld, M->d_cdense.p, M->d_ctile.p, M->d_cinv.p, s);
    M->cN = n;
    M->cLd = ld;
    M->coarseState = 1;
    return true;
}

static void mixed_pcg(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                      fh_perf* perf)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    for (int i = 0; i < 2; i++) M->v[i].alloc(N);
    double* rA = M->v[0].p;
    for (int i = 0; i < 3; i++) M->f[i].alloc(N);
    float *r32 = M->f[0].p, *p32 = M->f[1].p, *w32 = M->f[2].p;
    ensure_f32(M);
    int precon = c->preconditioner;
    if (c->innerPreconditioner >= 0) precon = c->innerPreconditioner;
    else if (M->mesh->tiled && (precon == FH_PRECON_DIC || precon == FH_PRECON_FDIC))
        precon = FH_PRECON_TILE_DIC;
    if (precon != FH_PRECON_NONE && factor_kind(precon) >= 0)
        (void)rD_for<float>(M, factor_kind(precon));

    state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
    init_fp64(M, psi, src, rA, M->v[1].p);
    launch_convert<double, float>(N, rA, r32, M->stream);
    // resLast = initial residual, replaceRatio: one decade per replacement
    launch_mixed_setup(st, 0.1, M->stream);
    const float* diag = M->d_diag32.p;
    const float* val = M->d_val32.p;
    // hex boxes: the FP32 product reads direction-packed coefficients
    const bool box = M->mesh->tiled;
    const float* dir[6] = {};
    // and the FP64 residual replacements the packed FP64 coefficients
    const double* dir64[6] = {};
    if (box)
    {
        ensure_dir32(M);
        ensure_dir64(M);
        for (int d = 0; d < 6; d++)
        {
            dir[d] = M->d_dir32[d].p;
            dir64[d] = M->d_dir64[d].p;
        }
    }
    const fh_mesh* mesh = M->mesh;
    // boxes with nx a multiple of 4: p = w + beta p fused into the product,
    // ping-ponging between two p buffers
    const bool fuseP = box && (mesh->boxNx & 3) == 0;
    M->f[3].alloc(N);
    M->f[4].alloc(N);
    float* pBuf[2] = {p32, M->f[3].p};
    // the FP32 solution correction since the last fold into psi
    M->f[5].alloc(N);
    float* xc = M->f[5].p;
    FH_HIP(hipMemsetAsync(xc, 0, (size_t)N*sizeof(float), M->stream));
    // A p: in place of w when p is formed by its own kernel; a separate
    // buffer when the product reads w to form p
    float* ap = fuseP ? M->f[4].p : w32;
    // two-level inner preconditioner on boxes (tile-local DIC plus the tile
    // coarse correction, carried into p by the fused product, or added to w
    // before p is formed)
    const bool coarse = precon == FH_PRECON_TILE_DIC && box && ensure_coarse(M);
    // rows not a multiple of 16: the tile-local apply reads r on padded
    // rows; the update and the replacement write that copy as well, so from
    // the second iteration on the apply need not pad r (the first pads it)
    const bool padR = box && (precon == FH_PRECON_TILE_DIC || precon == FH_PRECON_TILE_DILU)
                   && (mesh->boxNx & 15);
    const int32_t nxp = (mesh->boxNx + 15) & ~15;
    run_batches(M, c->maxIter, [&](int32_t seq)
    {
        M->padXFresh = padR && seq > 0;
        kt_mark(M, FH_KERNEL_PRECONDITION);
        if (coarse)
        {
            // two-level: the tile-local apply leaves its w.r partials and the
            // tile sums of r; the coarse solve adds P Ac^-1 P^T r (through
            // p in the product) and finalizes the whole w.r (the live timing
            // covers the tile-local apply only)
            // the non-fused product path reads w padded, where the apply pads
            M->padKeep = !fuseP;
            M->padLeft = false;
            const int32_t nbw = apply_precond<float>(M, precon, r32, w32, part, st, false,
                                                     FusedFinalize(), nullptr, M->d_crc.p);
            M->padKeep = false;
            kt_mark(M, FH_KERNEL_PRECONDITION);
            const FusedFinalize fc = fused_fin(M, FIN_PCG_WARA);
            double* partC = part + 2*(size_t)nbw;
            const int32_t nbc = launch_coarse_apply(M->cN, M->cLd, M->d_cinv.p, M->d_crc.p,
                                                    M->d_cz.p, part, nbw, partC, fc, st,
                                                    M->stream);
            if (!fc.counter)
                launch_finalize_blocks(nbc, N, partC, FIN_PCG_WARA, st, 0, M->stream);
        }
        else
        {
            precond_finalize<float>(M, precon, r32, w32, FIN_PCG_WARA);
            kt_mark(M, FH_KERNEL_PRECONDITION);
        }
        if (fuseP)
        {
            p32 = pBuf[(seq + 1) & 1];
            const FusedFinalize fa = fused_fin(M, FIN_PCG_WAPA);
            kt_mark(M, FH_KERNEL_SPMV);
            const int32_t nb = launch_box_pcg_spmv<float>(mesh->boxNx, mesh->boxNy, mesh->boxNz,
                                                       diag, dir, M->symmetric, w32,
                                                       pBuf[seq & 1], p32, ap, part, st,
                                                       M->stream, fa,
                                                       coarse ? M->d_cz.p : nullptr);
            kt_mark(M, FH_KERNEL_SPMV);
            if (!fa.counter)
                launch_finalize_blocks(nb, N, part, FIN_PCG_WAPA, st, 0, M->stream);
        }
        else if (box)
        {
            // p = (w + z) + beta p, as the fused product forms it
            if (M->padLeft)
                launch_pcg_p_padded(mesh->boxNx, mesh->boxNy, mesh->boxNz,
                                    (mesh->boxNx + 15) & ~15, M->d_padW.p,
                                    coarse ? M->d_cz.p : nullptr, p32, st, M->stream);
            else
            {
                if (coarse)
                    launch_add_tile_correction(mesh->boxNx, mesh->boxNy, mesh->boxNz, M->d_cz.p,
                                               w32, st, M->stream);
                launch_pcg_p<float>(N, w32, p32, st, M->stream);
            }
            M->padLeft = false;
            const int32_t nb = launch_box_spmv_dot_f32(mesh->boxNx, mesh->boxNy, mesh->boxNz,
                                                       diag, dir, M->symmetric, p32, w32, p32,
                                                       part, st, M->stream);
            launch_finalize_blocks(nb, N, part, FIN_PCG_WAPA, st, 0, M->stream);
        }
        else
        {
            launch_pcg_p<float>(N, w32, p32, st, M->stream);
            launch_spmv_dot<float>(M->rows(), diag, val, p32, w32, p32, nullptr, part, st,
                                   M->stream);
            launch_finalize(N, part, FIN_PCG_WAPA, st, 0, M->stream);
        }
        M->padXFresh = false;
        float* rPad = padR ? M->d_padX.p : nullptr;
        const FusedFinalize fu = fused_fin(M, FIN_MIXED_RES);
        kt_mark(M, FH_KERNEL_AXPY_NORM);
        const int32_t nbu = launch_mixed_update(N, psi, xc, r32, p32, ap, part, st, M->stream,
                                                fu, mesh->boxNx, nxp, rPad);
        kt_mark(M, FH_KERNEL_AXPY_NORM);
        if (!fu.counter) launch_finalize_blocks(nbu, N, part, FIN_MIXED_RES, st, 0, M->stream);
        const FusedFinalize ft = fused_fin(M, FIN_MIXED_TRUE);
        const int32_t nbr =
            box ? launch_mixed_replace_box(mesh->boxNx, mesh->boxNy, mesh->boxNz, M->d_diag.p,
                                           dir64, M->symmetric, psi, xc, src, r32, part, st,
                                           M->stream, ft, nxp, rPad)
                : launch_mixed_replace(M->rows(), M->d_diag.p, M->d_val.p, psi, xc, src, r32,
                                       part, st, M->stream, ft);
        if (!ft.counter) launch_finalize_blocks(nbr, N, part, FIN_MIXED_TRUE, st, 0, M->stream);
    });
    // the correction of the last iterations
    launch_axpy_f32(N, psi, xc, 1.0, M->stream);
    const DevState* h = M->h_state;
    std::memset(perf, 0, sizeof(*perf));
    std::snprintf(perf->solverName, sizeof(perf->solverName), "%sPCG(mixed)",
                  precon_name(c->preconditioner));
    perf->initialResidual = h->initialResidual;
    perf->finalResidual = h->finalResidual;
    perf->nIterations = h->nIterations;
    perf->converged = h->converged;
    perf->singular = h->singular;
    perf->nRefinements = h->nReplace;
    perf->normFactor = h->normFactor;
}

// ---------------------------------------------------------------------------
// mixed precision: FP64 outer residual / normFactor / convergence test /
// solution update around restarted FP32 inner solves (PBiCGStab)
// ---------------------------------------------------------------------------
static void mixed_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                        fh_perf* perf)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    for (int i = 0; i < 2; i++) M->v[i].alloc(N);
    double* rA = M->v[0].p;
    double* tmp = M->v[1].p;
    for (int i = 0; i < 9; i++) M->f[i].alloc(N);
    ensure_f32(M);

    state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
    init_fp64(M, psi, src, rA, tmp);
    FH_HIP(hipMemcpyAsync(M->h_state, M->d_state.p, sizeof(DevState), hipMemcpyDeviceToHost,
                          M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
    DevState outer = *M->h_state;
    int32_t totalInner = 0, nRefine = 0;
    const double innerRel = c->innerRelTol > 0 ? c->innerRelTol : 1e-4;
    const int32_t maxRef = c->maxRefinements > 0 ? c->maxRefinements : 50;
    bool singular = false;

    while (!outer.done && !outer.numericError)
    {
        if (nRefine >= maxRef || totalInner >= c->maxIter) break;
        // scaled FP32 right-hand side r / s, s = mean |r|
        const double s = outer.finalResidual*outer.normFactor/N;
        const double scale = (s > 0 && std::isfinite(s)) ? s : 1.0;
        launch_scale_convert(N, rA, 1.0/scale, M->f[0].p, M->stream);     // r32
        FH_HIP(hipMemsetAsync(M->f[1].p, 0, sizeof(float)*N, M->stream));  // d32 = 0
        // inner controls: relative reduction only
        const double target = std::max(innerRel, 0.3*c->tolerance/std::max(outer.finalResidual, 1e-300));
        state_upload(M, c, 0.0, std::min(target, 0.5), c->maxIter - totalInner, 0);
        launch_sum_abs<float>(N, M->f[0].p, part, M->stream);
        launch_finalize(N, part, FIN_INNER_INIT, st, 0, M->stream);
        float* r32 = M->f[0].p;
        float* d32 = M->f[1].p;
        if (c->solver == FH_SOLVER_PBICGSTAB)
        {
            float* V[7] = {M->f[2].p, M->f[3].p, M->f[4].p, M->f[5].p, M->f[6].p, M->f[7].p,
                           M->f[8].p};
            bicgstab_loop<float>(M, c->preconditioner, d32, r32, V, c->maxIter - totalInner);
        }
        else
        {
            pcg_loop<float>(M, c->preconditioner, d32, r32, M->f[2].p, M->f[3].p,
                            c->maxIter - totalInner);
        }
        const DevState inner = *M->h_state;
        totalInner += inner.nIterations;
        singular = inner.singular;
        // psi += scale*d; r = source - A psi; outer convergence test
        launch_axpy_f32(N, psi, d32, scale, M->stream);
        state_upload_outer(M, outer);
        launch_residual<double>(M->rows(), M->d_diag.p, M->d_val.p, psi, src, rA, nullptr,
                                M->stream);
        launch_sum_abs<double>(N, rA, part, M->stream);
        launch_finalize(N, part, FIN_OUTER_RES, st, 0, M->stream);
        FH_HIP(hipMemcpyAsync(M->h_state, M->d_state.p, sizeof(DevState),
                              hipMemcpyDeviceToHost, M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
        outer = *M->h_state;
        nRefine++;
        if (inner.nIterations == 0 && !outer.done) break;   // no progress possible
    }
    std::memset(perf, 0, sizeof(*perf));
    std::snprintf(perf->solverName, sizeof(perf->solverName), "%s%s(mixed)",
                  precon_name(c->preconditioner),
                  c->solver == FH_SOLVER_PBICGSTAB ? "PBiCGStab" : "PCG");
    perf->initialResidual = outer.initialResidual;
    perf->finalResidual = outer.finalResidual;
    perf->nIterations = totalInner;
    perf->converged = outer.converged;
    perf->singular = singular && !outer.converged;
    perf->nRefinements = nRefine;
    perf->normFactor = outer.normFactor;
    if (outer.numericError) fail(FH_ERR_NUMERIC, "non-finite residual in mixed-precision solve");
}

// ---------------------------------------------------------------------------
// PBiCG (PBiCG.C:65-222): the transposed system runs alongside with Tmul
// and preconditionT
// ---------------------------------------------------------------------------
static void pbicg_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                        fh_perf* perf)
{
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    hipStream_t s = M->stream;
    for (int i = 4; i < 10; i++) M->v[i].alloc(N);
    double *rA = M->v[4].p, *wA = M->v[5].p, *pA = M->v[6].p;
    double *rT = M->v[7].p, *wT = M->v[8].p, *pT = M->v[9].p;
    const int precon = c->preconditioner;
    state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
    init_fp64(M, psi, src, rA, wA);
    ensure_valT(M);
    const double* diag = M->d_diag.p;
    const double* val = M->d_val.p;
    const double* valT = M->d_valT.p;
    if (precon != FH_PRECON_NONE && factor_kind(precon) >= 0) (void)ensure_rD(M, factor_kind(precon));
    // rT = source - T psi
    launch_amul<double>(M->rows(), diag, valT, psi, wT, st, s);
    launch_diff_gamg<double>(N, src, wT, rT, st, s);
    run_batches(M, c->maxIter, [&](int32_t)
    {
        apply_precond<double>(M, precon, rA, wA, nullptr, st, false);
        apply_precond<double>(M, precon, rT, wT, nullptr, st, true);
        launch_dot<double, double>(N, wA, rT, part, st, s);
        launch_finalize(N, part, FIN_PCG_WARA, st, 0, s);
        launch_pcg_p<double>(N, wA, pA, st, s);
        launch_pcg_p<double>(N, wT, pT, st, s);
        launch_amul<double>(M->rows(), diag, valT, pT, wT, st, s);
        launch_spmv_dot<double>(M->rows(), diag, val, pA, wA, pT, nullptr, part, st, s);
        launch_finalize(N, part, FIN_PCG_WAPA, st, 0, s);
        launch_axpy_alpha<double>(N, rT, wT, st, s);
        launch_pcg_update<double>(N, psi, rA, pA, wA, part, st, s);
        launch_finalize(N, part, FIN_PCG_RES, st, 0, s);
    });
    char name[64];
    std::snprintf(name, sizeof(name), "%sPBiCG", precon_name(precon));
    fill_perf(M, perf, name);
}

// the controls a GAMG preconditioner is built from: its own sub-dictionary's
// when given (lduMatrixPreconditioner.C:72-82), else the solver's
static const fh_controls* gamg_precond_controls(const fh_controls* c)
{
    return c->preconditionerControls ? c->preconditionerControls : c;
}

// ---------------------------------------------------------------------------
void solve_device(fh_matrix* M, const fh_controls* c, double* d_psiNat,
                  const double* d_srcNat, fh_perf* perf)
{
    if (!M->hasCoeffs) fail(FH_ERR_INVALID, "matrix coefficients not set");
    validate_selection(M, c);
    if (M->mesh->nIfFaces > 0)
    {
        if (!M->ifSet)
            fail(FH_ERR_INVALID, "the mesh has coupled interfaces: set the matrix's interface "
                                 "coefficients (fh_matrix_set_interfaces) before solving");
        // GAMG would need the coupled interfaces agglomerated onto every
        // level (GAMGInterface); the mixed precision path has no interface
        // terms in its box kernels either way
        if (c->solver == FH_SOLVER_GAMG || c->preconditioner == FH_PRECON_GAMG)
            fail(FH_ERR_UNSUPPORTED, "GAMG with coupled interfaces is not supported");
    }
    const int32_t N = M->mesh->nCells;
    const int32_t* perm = M->mesh->d_perm.p;
    hipStream_t s = M->stream;
    for (int i = 0; i < 2; i++) M->v[i].alloc(N);
    M->v[2].alloc(N);
    M->v[3].alloc(N);
    double* psi = M->v[2].p;
    double* src = M->v[3].p;
    FH_HIP(hipEventRecord(M->evStart, s));
    launch_permute_in<double, double>(N, perm, d_psiNat, psi, s);
    launch_permute_in<double, double>(N, perm, d_srcNat, src, s);

    if (M->diagonalOnly || c->solver == FH_SOLVER_DIAGONAL)
    {
        // diagonalSolver.C:62-81
        launch_divide(N, src, M->d_diag.p, psi, s);
        std::memset(perf, 0, sizeof(*perf));
        std::snprintf(perf->solverName, sizeof(perf->solverName), "diagonal");
        perf->converged = 1;
    }
    else if (c->precision == FH_PRECISION_MIXED && c->solver == FH_SOLVER_PCG)
    {
        if (c->preconditioner == FH_PRECON_GAMG) gamg_prepare(M, gamg_precond_controls(c));
        mixed_pcg(M, c, psi, src, perf);
    }
    else if (c->precision == FH_PRECISION_MIXED && c->solver == FH_SOLVER_PBICGSTAB)
    {
        if (c->preconditioner == FH_PRECON_GAMG) gamg_prepare(M, gamg_precond_controls(c));
        mixed_solve(M, c, psi, src, perf);
    }
    else if (c->precision == FH_PRECISION_MIXED && c->solver == FH_SOLVER_GAMG)
    {
        gamg_solve(M, c, psi, src, perf, true);
    }
    else
    {
        if (c->preconditioner == FH_PRECON_GAMG
            && (c->solver == FH_SOLVER_PCG || c->solver == FH_SOLVER_PBICGSTAB
                || c->solver == FH_SOLVER_PBICG))
            gamg_prepare(M, gamg_precond_controls(c));
        switch (c->solver)
        {
            case FH_SOLVER_PCG:
            {
                for (int i = 4; i < 7; i++) M->v[i].alloc(N);
                double *rA = M->v[4].p, *pA = M->v[5].p, *wA = M->v[6].p;
                state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
                init_fp64(M, psi, src, rA, wA);
                pcg_loop<double>(M, c->preconditioner, psi, rA, pA, wA, c->maxIter);
                FH_HIP(hipMemcpyAsync(M->h_state, M->d_state.p, sizeof(DevState),
                                      hipMemcpyDeviceToHost, s));
                FH_HIP(hipStreamSynchronize(s));
                char name[64];
                std::snprintf(name, sizeof(name), "%sPCG", precon_name(c->preconditioner));
                fill_perf(M, perf, name);
                break;
            }
            case FH_SOLVER_PBICGSTAB:
            {
                for (int i = 4; i < 12 && i < 10; i++) M->v[i].alloc(N);
                double* rA = M->v[4].p;
                double* tmp = M->v[5].p;
                state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
                init_fp64(M, psi, src, rA, tmp);
                DBuf<double> extra[2];
                extra[0].alloc(N);
                extra[1].alloc(N);
                double* V[7] = {M->v[5].p, M->v[6].p, M->v[7].p, M->v[8].p, M->v[9].p,
                                extra[0].p, extra[1].p};
                FH_HIP(hipMemcpyAsync(M->h_state, M->d_state.p, sizeof(DevState),
                                      hipMemcpyDeviceToHost, s));
                FH_HIP(hipStreamSynchronize(s));
                if (!M->h_state->done)
                    bicgstab_loop<double>(M, c->preconditioner, psi, rA, V, c->maxIter);
                FH_HIP(hipMemcpyAsync(M->h_state, M->d_state.p, sizeof(DevState),
                                      hipMemcpyDeviceToHost, s));
                FH_HIP(hipStreamSynchronize(s));
                char name[64];
                std::snprintf(name, sizeof(name), "%sPBiCGStab", precon_name(c->preconditioner));
                fill_perf(M, perf, name);
                break;
            }
            case FH_SOLVER_PBICG:
            {
                pbicg_solve(M, c, psi, src, perf);
                break;
            }
            case FH_SOLVER_SMOOTH:
            {
                state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
                smooth_solver(M, c, psi, src);
                FH_HIP(hipStreamSynchronize(s));
                fill_perf(M, perf, "smoothSolver");
                break;
            }
            case FH_SOLVER_GAMG:
                gamg_solve(M, c, psi, src, perf, false);
                break;
            default:
                fail(FH_ERR_INVALID, "unknown solver id %d", c->solver);
        }
    }
    launch_permute_out<double, double>(N, perm, psi, d_psiNat, s);
    FH_HIP(hipEventRecord(M->evStop, s));
    FH_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    FH_HIP(hipEventElapsedTime(&ms, M->evStart, M->evStop));
    perf->solveMilliseconds = ms;
    check_sweep_errors(M);
    gamg_check_sweep_errors(M);
    if (M->h_state->scheduleError)
        fail(FH_ERR_SCHEDULE, "a sweep dependency wait expired");
    if (M->h_state->numericError)
        fail(FH_ERR_NUMERIC, "non-finite residual");
}

// ---------------------------------------------------------------------------
// measurement: back-to-back launches of one kernel between HIP events
// ---------------------------------------------------------------------------
template<class T>
static void profile_t(fh_matrix* M, int kernel, int precon, int reps, double* ms,
                      double* bytes)
{
    const int32_t N = M->mesh->nCells;
    const int64_t F = M->mesh->nFaces;
    const double w = sizeof(T);
    DBuf<T> x, y, z;
    x.alloc(N); y.alloc(N); z.alloc(N);
    FH_HIP(hipMemsetAsync(x.p, 0, x.bytes(), M->stream));
    FH_HIP(hipMemsetAsync(y.p, 0, y.bytes(), M->stream));
    FH_HIP(hipMemsetAsync(z.p, 0, z.bytes(), M->stream));
    const T* diag = diag_for<T>(M);
    const T* val = val_for<T>(M, false);
    const int which = factor_kind(precon);
    if ((kernel == FH_KERNEL_SWEEP_FORWARD || kernel == FH_KERNEL_SWEEP_BACKWARD
         || kernel == FH_KERNEL_PRECONDITION) && (which < 0 || which == 2))
        fail(FH_ERR_INVALID, "sweep profiling needs DIC, FDIC, DILU, tileDIC or tileDILU");
    const T* rD = which >= 0 ? rD_for<T>(M, which) : nullptr;
    const bool local = which >= 3 && M->mesh->tiled;
    const fh_mesh* mesh = M->mesh;
    // boxes with nx a multiple of 4 run the fused PCG product and the
    // packed-coefficient sweeps with their fused w.r sums (the kernels of
    // the PCG solves); other meshes the general kernels
    const bool box = mesh->tiled;
    const bool boxX4 = box && (mesh->boxNx & 3) == 0;
    const T* dir[6] = {};
    if (box)
    {
        if constexpr (std::is_same<T, float>::value)
        {
            ensure_dir32(M);
            for (int q = 0; q < 6; q++) dir[q] = M->d_dir32[q].p;
        }
        else
        {
            ensure_dir64(M);
            for (int q = 0; q < 6; q++) dir[q] = M->d_dir64[q].p;
        }
    }
    // the box product reads a solver state: a steady-state iteration
    // (p = w + beta*pOld), never done
    DBuf<DevState> pst;
    pst.alloc(1);
    DevState hs{};
    hs.nIterations = 1;
    hs.beta = 0.5;
    FH_HIP(hipMemcpyAsync(pst.p, &hs, sizeof(DevState), hipMemcpyHostToDevice, M->stream));
    const DevState* pstate = pst.p;
    DBuf<T> p2;
    if (boxX4) { p2.alloc(N); FH_HIP(hipMemsetAsync(p2.p, 0, p2.bytes(), M->stream)); }
    DBuf<double> psi64;
    if (std::is_same<T, float>::value && kernel == FH_KERNEL_AXPY_NORM)
    {
        psi64.alloc(N);
        FH_HIP(hipMemsetAsync(psi64.p, 0, psi64.bytes(), M->stream));
    }
    SweepData<T> d{};
    d.val = val; d.rD = rD; d.x = x.p; d.w = y.p;
    double* part = M->d_partials.p;
    auto once = [&]()
    {
        switch (kernel)
        {
            case FH_KERNEL_SPMV:
                if (boxX4)
                {
                    launch_box_pcg_spmv<T>(mesh->boxNx, mesh->boxNy, mesh->boxNz, diag, dir,
                                           M->symmetric, x.p, z.p, p2.p, y.p, part, pstate,
                                           M->stream);
                    break;
                }
                launch_amul<T>(M->rows(), diag, val, x.p, y.p, nullptr, M->stream);
                break;
            case FH_KERNEL_SWEEP_FORWARD:
            case FH_KERNEL_SWEEP_BACKWARD:
            {
                const bool bwd = kernel == FH_KERNEL_SWEEP_BACKWARD;
                if constexpr (std::is_same<T, float>::value)
                {
                    if (local)
                    {
                        // the kernels apply_precond runs for FP32 tile-local factors
                        const TileView G = M->tile_view(TileBZ<float>::pre);
                        if (bwd)
                            launch_tile_local_f32(true, G, y.p, rD, M->d_dir32[5].p,
                                                  M->d_dir32[4].p, M->d_dir32[3].p, y.p,
                                                  nullptr, M->stream, x.p, part);
                        else
                            launch_tile_local_f32(false, G, x.p, rD, M->d_dir32[0].p,
                                                  M->d_dir32[1].p, M->d_dir32[2].p, y.p,
                                                  nullptr, M->stream);
                        break;
                    }
                }
                if (box && !local)
                {
                    box_precond_pass<T>(M, bwd, d, bwd ? part : nullptr, nullptr);
                    break;
                }
                sweep_t<T>(M, bwd ? SWEEP_PRECOND_BWD : SWEEP_PRECOND_FWD, d, nullptr, local);
                break;
            }
            case FH_KERNEL_PRECONDITION:
                apply_precond<T>(M, precon, x.p, y.p, part, pst.p, false);
                break;
            case FH_KERNEL_AXPY_NORM:
                if constexpr (std::is_same<T, float>::value)
                {
                    // the update of the mixed inner iteration: FP64 solution,
                    // FP32 recurrence residual
                    launch_mixed_update(N, psi64.p, x.p, y.p, z.p, z.p, part, pstate,
                                        M->stream);
                    break;
                }
                launch_pcg_update_nostate<T>(N, x.p, y.p, z.p, z.p, M->d_partials.p, M->stream);
                break;
            case FH_KERNEL_H:
                if constexpr (std::is_same<T, float>::value)
                    fail(FH_ERR_INVALID, "H is profiled in FP64 only");
                else
                    launch_h(M->rows(), M->d_val.p, 1, x.p, y.p, M->stream);
                break;
            default:
                fail(FH_ERR_INVALID, "unknown kernel id %d", kernel);
        }
    };
    once();
    FH_HIP(hipStreamSynchronize(M->stream));
    FH_HIP(hipEventRecord(M->evStart, M->stream));
    for (int i = 0; i < reps; i++) once();
    FH_HIP(hipEventRecord(M->evStop, M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
    float t = 0.f;
    FH_HIP(hipEventElapsedTime(&t, M->evStart, M->evStop));
    *ms = (double)t/reps;
    switch (kernel)
    {
        case FH_KERNEL_SPMV:
            // boxes: the fused PCG product reads w, pOld, diag and one
            // coefficient per face (the symmetric upper one from the lower
            // array of the neighbour) and writes pNew and y; otherwise
            // diag, x, y, row info and (value, column) per off-diagonal
            if (boxX4)
                *bytes = N*5*w + (M->symmetric ? 1.0 : 2.0)*F*w;
            else
                *bytes = N*(3*w + 4) + 2.0*F*(w + 4);
            break;
        case FH_KERNEL_SWEEP_FORWARD:
        case FH_KERNEL_SWEEP_BACKWARD:
            // boxes (packed coefficients): input, rD, output and one
            // coefficient per face, plus r for the fused w.r sum of the
            // backward pass; otherwise input, rD, output, row info and
            // (value, column) of one triangle
            if (box)
                *bytes = N*(kernel == FH_KERNEL_SWEEP_BACKWARD ? 4 : 3)*w + 1.0*F*w;
            else
                *bytes = N*(3*w + 4) + 1.0*F*(w + 4);
            break;
        case FH_KERNEL_PRECONDITION:
            if (box && local && w == 4 && (mesh->boxNx & 15) == 0)
            {
                // one launch: x, rD, one coefficient per face (each triangle
                // from the lower arrays when symmetric; FP16 for the
                // symmetric tile-local DIC), w written once
                const double cw = which == 3 && tile_dic_half(M) ? 2.0 : (double)w;
                *bytes = N*3*w + (M->symmetric ? 1.0 : 2.0)*F*cw;
            }
            else if (box && !local)
                // two passes: x, rD, w, then w, rD, w, r; a coefficient
                // per face in each
                *bytes = N*7*w + 2.0*F*w;
            else
                *bytes = 2.0*(N*(3*w + 4) + 1.0*F*(w + 4)) + N*w;
            break;
        case FH_KERNEL_AXPY_NORM:
            // FP32 (mixed): psi FP64 read+write, r read+write, p, w read;
            // FP64: psi, r read+write, p, w read
            // FP32 (the mixed update between folds): correction and
            // recurrence residual read and written, p and A p read
            *bytes = 6.0*N*w; break;
        case FH_KERNEL_H:
            // x, y and row info per cell; (value, column) per off-diagonal
            *bytes = N*(2*w + 4) + 2.0*F*(w + 4); break;
    }
}

void profile_kernel(fh_matrix* M, int kernel, int precision, int precon, int reps,
                    double* ms, double* bytes)
{
    if (reps <= 0) fail(FH_ERR_INVALID, "reps must be positive");
    if (precision == FH_PRECISION_MIXED) profile_t<float>(M, kernel, precon, reps, ms, bytes);
    else profile_t<double>(M, kernel, precon, reps, ms, bytes);
}

} // namespace fh
