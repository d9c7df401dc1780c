// Internal definition of fh_matrix: the device-resident coefficients of one
// lduMatrix on a mesh, its cached preconditioner factors and the work space
// of the solvers that run on it.
#pragma once

#include "fh_kernels.h"
#include "fh_solver_kernels.h"

struct fh_matrix
{
    fh_mesh* mesh = nullptr;
    hipStream_t stream = nullptr;
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

    // FP32 copies for the mixed-precision inner solves
    fh::DBuf<float> d_diag32, d_val32;
    bool f32Valid = false;

    // preconditioner / smoother factors (library numbering)
    fh::DBuf<double> d_rD_DIC, d_rD_DILU, d_rD_diag;
    bool rDValid[3] = {false, false, false};
    fh::DBuf<float> d_rD32_DIC, d_rD32_DILU, d_rD32_diag;
    bool rD32Valid[3] = {false, false, false};

    // pipelined sweep synchronisation
    fh::DBuf<unsigned long long> d_sync;  // [0] ticket, [1..nParts] progress
    unsigned long long epoch = 0;

    // device scalar state and reduction partials
    fh::DBuf<fh::DevState> d_state;
    fh::DevState* h_state = nullptr;      // pinned
    fh::DBuf<double> d_partials;

    // work vectors (library numbering)
    fh::DBuf<double> v[10];
    fh::DBuf<float> f[10];
    fh::DBuf<double> d_psiNat, d_srcNat;  // staging for host-array solves

    fh::RowsView rows() const
    {
        fh::RowsView R;
        R.n = mesh->nCells;
        R.rowInfo = mesh->d_rowInfo.p;
        R.sliceStart = mesh->d_sliceStart.p;
        R.col = mesh->d_col.p;
        R.face = mesh->d_face.p;
        return R;
    }

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
void matrix_init(fh_matrix* M, fh_mesh* mesh);
void matrix_free(fh_matrix* M);
void matrix_set(fh_matrix* M, const double* diag, const double* upper, const double* lower,
                bool device);
void matrix_coeffs_updated(fh_matrix* M);
void ensure_sumA(fh_matrix* M);
void ensure_valT(fh_matrix* M);
void ensure_f32(fh_matrix* M);
// factor of the DIC/FDIC (0), DILU (1) or diagonal (2) kind, FP64 or FP32
const double* ensure_rD(fh_matrix* M, int which);
const float* ensure_rD32(fh_matrix* M, int which);

// preconditioner apply w = M^-1 r on library-numbered vectors.  For the
// FP64 PCG path a fused dot (w.r) can be requested (partials != null).
template<class T>
void apply_precond(fh_matrix* M, int precon, const T* r, T* w, double* partials,
                   DevState* st, bool transpose);

void profile_kernel(fh_matrix* M, int kernel, int precision, int precon, int reps,
                    double* ms, double* bytes);
void solve_device(fh_matrix* M, const fh_controls* c, double* d_psiNat,
                  const double* d_srcNat, fh_perf* perf);
void smooth_device(fh_matrix* M, int smoother, double* psi, const double* src,
                   int32_t nSweeps, DevState* st);
} // namespace fh
