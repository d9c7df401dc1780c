// Launch wrappers of the HIP kernels (fh_kernels.hip).
#pragma once

#include "fh_internal.h"

namespace fh
{

// Device-resident scalar state of one iterative solve.  The reduction
// finalize kernels update it, every iteration kernel reads `done` first,
// so the host enqueues iterations in batches without a sync per iteration.
struct DevState
{
    double normFactor;
    double initialResidual;
    double finalResidual;
    double tolerance;
    double relTol;
    // PCG
    double wArA, wArAold, wApA, alpha, beta;
    // PBiCGStab
    double rA0rA, rA0rAold, rA0AyA, omega, tAtA, tAsA, bicgAlpha, bicgBeta;
    // scratch sums
    double sum0, sum1;
    int32_t maxIter, minIter;
    int32_t nIterations;
    int32_t converged;
    int32_t singular;
    int32_t done;
    int32_t numericError;
    int32_t scheduleError;
    int32_t earlySeq;       // PBiCGStab early exit: launch sequence number
    // mixed precision reliable updates
    int32_t replace;        // 1: recompute the FP64 residual this iteration
    int32_t nReplace;       // FP64 residual replacements made
    int32_t foldPending;    // 1: the next update folds the FP32 correction into x
    double resLast;         // true residual at the last replacement
    double replaceRatio;    // replace when the FP32 residual fell by this
    double gamgScale;       // GAMG correction scaling factor
};

// Row structure in the new numbering (shared by every kernel)
struct RowsView
{
    int32_t n;
    const int32_t* rowInfo;     // nLower | nEnt << 16
    const int32_t* sliceStart;  // SELL-64 slice offsets
    const int32_t* col;
    const int32_t* face;        // (f << 1 | isUpper), -1 pad
    // coupled interfaces (null: none): row r's interface entries are
    // ifStart[r] .. ifStart[r+1]-1, column ifNbr[e], value val[ifBase + e]
    // (the negated interface coefficient), after the row's face entries --
    // the reference adds the interfaces after the face loop
    const int32_t* ifStart = nullptr;
    const int32_t* ifNbr = nullptr;
    int64_t ifBase = 0;
};

// acc += sum of the row's interface entries val*x(nbr), in entry order
template<class T, class X>
__device__ __forceinline__ T add_interfaces(const RowsView& R, int32_t r, const T* val,
                                            const X* x, T acc)
{
    if (R.ifStart)
        for (int32_t e = R.ifStart[r]; e < R.ifStart[r + 1]; e++)
            acc += val[R.ifBase + e]*(T)x[R.ifNbr[e]];
    return acc;
}

// Partitioned level sweep schedule (one workgroup per part)
struct SchedView
{
    int32_t nParts;
    int32_t nLevels;
    const int32_t* partLevelOff;
    const int32_t* partMinLevel;
    const int32_t* segStart;
    const int32_t* upOff;
    const int32_t* upList;
    const int32_t* downOff;
    const int32_t* downList;
    const int32_t* order;
    int32_t maxDeps;
};

// Per-matrix synchronisation context of the pipelined sweeps
struct SweepSync
{
    unsigned long long* ticket;    // [1]
    unsigned long long* progress;  // [nParts]
    unsigned long long epoch;      // launches made with this context so far
};

enum SweepKind
{
    SWEEP_PRECOND_FWD = 0,  // w = rD*x, w -= rD*lower*w(l)      (DIC/DILU)
    SWEEP_PRECOND_BWD = 1,  // w -= rD*upper*w(u), faces reversed
    SWEEP_RD_DIC = 2,       // rD -= upper*upper/rD(l)
    SWEEP_RD_DILU = 3,      // rD -= upper*lower/rD(l)
    SWEEP_GS_FWD = 4,       // Gauss-Seidel forward row update
    SWEEP_GS_BWD = 5,       // symmetric Gauss-Seidel backward row update
    SWEEP_PRECOND_BWD_T = 6 // preconditionT backward: upper entries by
                            // descending neighbour (losort order reversed)
};

template<class T>
struct SweepData
{
    const T* val;        // SELL coefficients
    const T* rD;         // reciprocal factor (precond) / raw factor (RD)
    const T* x;          // input: residual (precond) or source (GS)
    T* w;                // output: w (precond), rD raw (RD), psi (GS)
    const T* diag;       // GS / RD
    const double* upperF;// RD_DILU: face-ordered upper
    const int32_t* perm; // PRECOND_BWD_T: library row -> original cell
    // hex boxes, PRECOND kinds: the pass's coefficients packed per cell by
    // direction (k, j, i), or null for the SELL values; dirShift: read the
    // coefficient at the neighbour (k+1, j+1, i+1), i.e. the upper
    // coefficient of a symmetric matrix from its lower-packed arrays
    const T* dir[3];
    int32_t dirShift;
    // PRECOND_BWD on boxes: also sum w*dotR into one partial per tile
    const T* dotR;
    double* partials;
};

// Chunks of the sync-free sweep: rows of one level, at most 64 per chunk,
// with the chunks each one reads in the forward / backward direction
struct ChunkView
{
    int32_t nChunks;
    const int32_t* start;      // [nChunks+1] library rows
    const int32_t* fwdOff;     // [nChunks+1]
    const int32_t* fwdDeps;    // chunks holding lower neighbours
    const int32_t* bwdOff;
    const int32_t* bwdDeps;    // chunks holding upper neighbours
};

// Synchronisation context of the chunk sweep
struct RowSync
{
    unsigned int* flags;               // per chunk: launch epoch + 1 when done
    unsigned long long* ticket;        // monotonic chunk ticket counter
    unsigned long long ticketBase;     // tickets handed out by earlier launches
    unsigned long long epoch;          // launches made with this context
    unsigned long long waitTicks;      // wall-clock bound of one wait
    unsigned int* err;                 // set when a wait expired
};

// persistent workgroups (4 wavefronts each) of one chunk-sweep launch
constexpr int32_t kRowSweepBlocks = 256;

// ---- kernels: sizes in rows --------------------------------------------------
template<class T>
void launch_gather_values(const RowsView& R, int64_t nnzPad, const double* upperF,
                          const double* lowerF, T* val, hipStream_t s);
template<class Tin, class Tout>
void launch_permute_in(int32_t n, const int32_t* perm, const Tin* xNat, Tout* xNew,
                       hipStream_t s);
template<class Tin, class Tout>
void launch_permute_out(int32_t n, const int32_t* perm, const Tin* xNew, Tout* xNat,
                        hipStream_t s);
template<class Tin, class Tout>
void launch_convert(int64_t n, const Tin* a, Tout* b, hipStream_t s);

// y = A x (T = float or double)
template<class T>
void launch_amul(const RowsView& R, const T* diag, const T* val, const T* x, T* y,
                 const DevState* st, hipStream_t s);
// y = A^T x
template<class T>
void launch_tmul(const RowsView& R, const T* diag, const T* valT, const T* x, T* y,
                 hipStream_t s);
// r = b - A x
template<class T>
void launch_residual(const RowsView& R, const T* diag, const T* val, const T* x,
                     const T* b, T* r, const DevState* st, hipStream_t s);
void launch_sum_a(const RowsView& R, const double* diag, const double* val,
                  double* sumA, hipStream_t s);
// bPrime = src - (interface entries)*psi for the Gauss-Seidel smoothers of
// a coupled mesh
void launch_interface_bprime(const RowsView& R, const double* val, const double* psi,
                             const double* src, double* bPrime, hipStream_t s);
// tail[e] = -coeffs[face[e]]: the interface entries' values
void launch_interface_values(int32_t nE, const int32_t* face, const double* coeffs,
                             double* tail, hipStream_t s);

// sweeps
// streaming probes: 0 read buf[0]; 1 copy buf[0] -> buf[1]; 2 buf[0] + buf[2]
// + buf[3] -> buf[1] (n4 16-byte elements each)
void launch_stream_probe(int pattern, int64_t n4, int grid, float4* const buf[4], float* out,
                         hipStream_t s);
// One colour of a multi-colour Gauss-Seidel sweep over the listed rows
template<class T>
void launch_sweep_colour(const RowsView& R, const SweepData<T>& d, const int32_t* rows,
                         int32_t n, const DevState* st, hipStream_t s);
// nSweeps multi-colour sweeps of a small mesh in one workgroup (colours
// ascending, then descending when sym; cstart: device copy of the colour
// offsets)
template<class T>
void launch_sweep_colours_wg(const RowsView& R, const SweepData<T>& d, const int32_t* rows,
                             const int32_t* cstart, int32_t nCol, int32_t nSweeps, bool sym,
                             const DevState* st, hipStream_t s);
// colours [k0, k1) (each of at most a few thousand rows) in one workgroup,
// ascending or descending
template<class T>
void launch_sweep_colour_run(const RowsView& R, const SweepData<T>& d, const int32_t* rows,
                             const int32_t* cstart, int32_t k0, int32_t k1, bool desc,
                             const DevState* st, hipStream_t s);
// One-workgroup pipelined level sweep of a one-part mesh (Gauss-Seidel and
// the DIC/DILU passes; fh_level_sweep.hip); false: kind or mesh not covered
template<class T>
bool launch_sweep_level_wg(int kind, const RowsView& R, const SchedView& S,
                           const SweepData<T>& d, DevState* st, hipStream_t s);
template<class T>
void launch_sweep(int kind, const RowsView& R, const SchedView& S, SweepSync& sync,
                  const SweepData<T>& d, DevState* st, hipStream_t s);
// tmp (nRows values, or null): the Gauss-Seidel and DIC/DILU passes poll
// dependency values in a sentinel-filled output (tmp for in-place passes)
// instead of waiting on chunk flags
template<class T>
void launch_sweep_rows(int kind, const RowsView& R, const SweepData<T>& d, const ChunkView& C,
                       RowSync& sync, DevState* st, hipStream_t s, int32_t nRows, T* tmp);
template<class T>
void launch_reciprocal(int32_t n, T* a, hipStream_t s);

// reductions (deterministic two-pass).  partials must hold
// reduction_partials(n) doubles per sum.
int32_t reduction_blocks(int32_t n);
size_t reduction_partials(int32_t n);

} // namespace fh
