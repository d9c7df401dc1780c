// GAMG device kernels (fh_gamg_kernels.hip).
#pragma once

#include "fh_kernels.h"

namespace fh
{

// Galerkin gathers of one agglomeration level: every coarse value is a sum
// over its fine contributions in the order the reference adds them
struct GalerkinMaps
{
    const int32_t *dStart, *dCells;   // coarse library row -> fine library rows
    const int32_t *iStart, *iFaces;   // coarse library row -> intra-cluster fine faces
    const int32_t *fStart, *fFaces;   // coarse face -> fine faces
    const uint8_t* fFlip;             // per gathered fine face: flipped
};

// The coarsest level in natural numbering for the single-workgroup Krylov
// solve (PCG+DIC when symmetric, PBiCGStab+DILU otherwise).
struct CoarsestSys
{
    int32_t n, nF;
    const int32_t* lower;      // natural face addressing
    const int32_t* upper;
    const int32_t* rowStart;   // natural rows: lower entries, then upper entries
    const int32_t* ent;        // f << 1 | isUpper
    const int32_t* colN;       // natural column of each entry
    const int32_t* perm;       // library row -> natural cell
    const double* diagLib;     // diag in library order
    const double* upperF;      // face order
    const double* lowerF;      // face order (== upperF when symmetric)
    double* diagN;             // natural copies written by the setup kernel
    double* rD;
    double* sumA;
    double* work;              // [13 n] when the vectors do not fit in LDS
    double tol, relTol;
    int32_t maxIter, minIter;
    int32_t sym;
};

template<class T>
void launch_restrict(int32_t nc, const int32_t* start, const int32_t* cells, const T* ff,
                     T* cf, const DevState* st, hipStream_t s);
template<class T>
void launch_prolong(int32_t nf, const int32_t* map, const T* cf, T* ff, const DevState* st,
                    hipStream_t s);
void launch_galerkin(int32_t nc, int32_t ncf, const GalerkinMaps& G, const double* fineDiagLib,
                     const double* fineUpper, const double* fineLower, double* coarseDiagLib,
                     double* cu, double* cl, hipStream_t s);
// field = sf*field + (source - sf*A field)/diag with sf from two sums
template<class T>
void launch_gamg_scale(const RowsView& R, const T* diag, const T* val, T* field, T* Acf,
                       const T* source, double* partials, DevState* st, hipStream_t s);

// coarsest level: iterative single-workgroup solve
void launch_coarsest_setup(const CoarsestSys& S, hipStream_t s);
template<class T>
void launch_coarsest_krylov(const CoarsestSys& S, const T* srcLib, T* corrLib,
                            const DevState* st, hipStream_t s);
// coarsest level: LU back-substitution (natural order, pivots from the host)
template<class T>
void launch_lu_backsub(int32_t n, const double* LU, const int32_t* pivot, const int32_t* perm,
                       double* work, const T* srcLib, T* corrLib, const DevState* st,
                       hipStream_t s);
// coarsest level: the dense matrix (library order, column-major)
void launch_dense_fill(int32_t n, const RowsView& R, const double* diag, const double* val,
                       double* A, hipStream_t s);
// coarsest level, n <= 96: Crout LU with implicit pivoting in the reference's
// order (the dense column-major library-order matrix in, natural order out)
void launch_crout_lu(int32_t n, const double* Acm, const int32_t* perm, double* LU,
                     int32_t* piv, int32_t* bad, hipStream_t s);
// coarsest level, larger: explicit inverse by Newton-Schulz X <- X (2I - A X)
// from X0 = A^T/(|A|_1 |A|_inf), every step on the device (FP64 MFMA
// products, the convergence test in the state); the host polls `done`
struct NsState
{
    double scale, prev, err;
    int32_t done, bad, cur, its;
};
// sums: 2n doubles
void launch_ns_init(int32_t n, const double* A, double* X0, double* sums, NsState* ns,
                    hipStream_t s);
void launch_ns_iterations(int32_t n, const double* A, double* X0, double* X1, double* T_,
                          double* partMax, NsState* ns, int iterations, hipStream_t s);
void launch_ns_final(int32_t n, const double* X1, double* X0, const NsState* ns, hipStream_t s);
// part: gemv_work_doubles(n) doubles
template<class T>
void launch_gemv(int32_t n, const double* X, const T* b, T* x, double* part,
                 const DevState* st, hipStream_t s);
int64_t gemv_work_doubles(int32_t n);

// small gated vector helpers
template<class T>
void launch_axpy_gamg(int32_t n, T* y, const T* x, const DevState* st, hipStream_t s);
template<class T>
void launch_sub_gamg(int32_t n, T* y, const T* x, const DevState* st, hipStream_t s);
template<class T>
void launch_diff_gamg(int32_t n, const T* a, const T* b, T* out, const DevState* st,
                      hipStream_t s);
template<class T>
void launch_copy_gamg(int32_t n, const T* a, T* b, const DevState* st, hipStream_t s);
template<class T>
void launch_zero_gamg(int32_t n, T* a, const DevState* st, hipStream_t s);
// mixed GAMG: b = (float)(a/s), then y += s*(double)x, s = mean |r| taken
// from the state's last residual sum
void launch_to_f32_gamg(int32_t n, const double* a, float* b, const DevState* st,
                        hipStream_t s);
void launch_add_f32_gamg(int32_t n, double* y, const float* x, const DevState* st,
                         hipStream_t s);

} // namespace fh
