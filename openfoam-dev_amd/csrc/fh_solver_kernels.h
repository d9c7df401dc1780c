// Launchers of the fused solver kernels (fh_solver_kernels.hip).
#pragma once

#include "fh_kernels.h"

namespace fh
{

enum FinalizeOp
{
    FIN_SUM = 0,
    FIN_XREF = 1,
    FIN_INIT = 2,
    FIN_PCG_WARA = 3,
    FIN_PCG_WAPA = 4,
    FIN_PCG_RES = 5,
    FIN_BICG_START = 6,
    FIN_BICG_ALPHA = 7,
    FIN_BICG_S = 8,
    FIN_BICG_OMEGA = 9,
    FIN_BICG_RES = 10,
    FIN_INNER_INIT = 11,
    FIN_OUTER_RES = 12,
    FIN_SMOOTH_RES = 13,
    FIN_MIXED_RES = 14,
    FIN_MIXED_TRUE = 15
};

// A reduction kernel that finalizes itself: the last of its blocks sums the
// partials and applies op to st.  counter: fin_counters(nb) zeroed device
// counters that the kernel leaves zeroed, and the partials buffer holds
// fin_partials(nb) values; null counter: the kernel only writes its
// partials and a separate finalize follows.
// each counter on a 128-byte line of its own: counters sharing a line
// serialize every block's arrival at that line (12 us for 1024 blocks)
constexpr int FIN_COUNTER_STRIDE = 32;
inline size_t fin_counters(size_t nb) { return FIN_COUNTER_STRIDE*(1 + (nb + 63)/64); }
inline size_t fin_partials(size_t nb) { return 2*nb + 2*((nb + 63)/64); }
struct FusedFinalize
{
    unsigned int* counter = nullptr;
    DevState* st = nullptr;
    int32_t op = 0;
    int32_t n = 0;
};
void launch_finalize(int32_t n, const double* partials, int32_t op, DevState* st,
                     int32_t seq, hipStream_t s);
template<class Ta, class Tb>
void launch_dot(int32_t n, const Ta* a, const Tb* b, double* partials, const DevState* st,
                hipStream_t s);
template<class T>
void launch_sum_abs(int32_t n, const T* a, double* partials, hipStream_t s);
void launch_init_residual(int32_t n, const double* Apsi, const double* source,
                          const double* sumA, double* rA, double* partials,
                          const DevState* st, hipStream_t s);
template<class T>
void launch_precond_point(int32_t n, int32_t kind, const T* rD, const T* r, T* w,
                          double* partials, const DevState* st, hipStream_t s);
template<class T>
void launch_pcg_p(int32_t n, const T* w, T* p, const DevState* st, hipStream_t s);
template<class T>
void launch_spmv_dot(const RowsView& R, const T* diag, const T* val, const T* x, T* y,
                     const T* z, const T* v, double* partials, const DevState* st,
                     hipStream_t s);
// y = A x on a hex box from direction-packed FP32 coefficients dir[0..5] =
// lower K, J, I, upper I, J, K (upper unused when symmetric), fused with
// sum y*z (z null: y*y).  Returns the number of partials it wrote (at most
// box_spmv_blocks), to be finalized with launch_finalize_blocks.
int32_t box_spmv_blocks(int32_t nx, int32_t ny, int32_t nz);
int32_t launch_box_spmv_dot_f32(int32_t nx, int32_t ny, int32_t nz, const float* diag,
                                const float* const dir[6], bool symmetric, const float* x,
                                float* y, const float* z, double* partials,
                                const DevState* st, hipStream_t s);
// PCG's p = w + beta*pOld (first iteration: w) fused into the box product
// (FP32 or FP64 direction-packed coefficients): writes pNew and y = A pNew
// with sum y*pNew; nx must be a multiple of 4 (returns -1 otherwise,
// nothing launched)
template<class T>
int32_t launch_box_pcg_spmv(int32_t nx, int32_t ny, int32_t nz, const T* diag,
                            const T* const dir[6], bool symmetric, const T* w,
                            const T* pOld, T* pNew, T* y, double* partials,
                            const DevState* st, hipStream_t s,
                            const FusedFinalize& F = FusedFinalize(), const T* zc = nullptr);
// finalize over an explicit number of partial blocks
void launch_finalize_blocks(int32_t nBlocks, int32_t n, const double* partials, int32_t op,
                            DevState* st, int32_t seq, hipStream_t s);
template<class T>
void launch_pcg_update(int32_t n, T* psi, T* r, const T* p, const T* w, double* partials,
                       const DevState* st, hipStream_t s,
                       const FusedFinalize& F = FusedFinalize());
template<class T>
void launch_bicg_p(int32_t n, const T* r, T* p, const T* AyA, const DevState* st,
                   hipStream_t s);
template<class T>
void launch_bicg_s(int32_t n, const T* r, const T* AyA, T* sA, double* partials,
                   const DevState* st, hipStream_t s);
template<class T>
void launch_bicg_early(int32_t n, T* psi, const T* y, const DevState* st, int32_t seq,
                       hipStream_t s);
template<class T>
void launch_bicg_update(int32_t n, T* psi, const T* y, const T* z, T* r, const T* sA,
                        const T* tA, const T* r0, double* partials, const DevState* st,
                        hipStream_t s);

template<class T>
void launch_pcg_update_nostate(int32_t n, T* psi, T* r, const T* p, const T* w,
                               double* partials, hipStream_t s);
// mixed PCG: x64 += alpha*p32, r32 -= alpha*w32, sum |r32|; returns the
// partial block count
// rPad (or null): r also written on rows of nx cells padded to nxp
int32_t launch_mixed_update(int32_t n, double* x, float* xc, float* r, const float* p,
                            const float* w, double* partials, const DevState* st,
                            hipStream_t s, const FusedFinalize& F = FusedFinalize(),
                            int32_t nx = 1, int32_t nxp = 1, float* rPad = nullptr);
// mixed PCG residual replacement (only when st->replace): r = b - A (x + xc)
// in FP64, r32 = r, sum |r|; returns the partial count
// the same on a box from the packed FP64 coefficients (bitwise equal)
int32_t launch_mixed_replace_box(int32_t nx, int32_t ny, int32_t nz, const double* diag,
                                 const double* const dir[6], bool symmetric, const double* x,
                                 const float* xc, const double* b, float* r32,
                                 double* partials, const DevState* st, hipStream_t s,
                                 const FusedFinalize& F, int32_t nxp = 1,
                                 float* rPad = nullptr);
int32_t launch_mixed_replace(const RowsView& R, const double* diag, const double* val,
                             const double* x, const float* xc, const double* b, float* r32,
                             double* partials, const DevState* st, hipStream_t s,
                             const FusedFinalize& F = FusedFinalize());
void launch_mixed_setup(DevState* st, double replaceRatio, hipStream_t s);
// psi += w
void launch_axpy_inplace(int32_t n, double* psi, const double* w, const DevState* st,
                         hipStream_t s);
// sum |x| (and sum x) with the done test
template<class T>
void launch_sum_abs_state(int32_t n, const T* a, double* partials, const DevState* st,
                          hipStream_t s);
void launch_finalize_smooth(int32_t n, const double* partials, DevState* st,
                            int32_t nSweeps, hipStream_t s);
// y32 = (float)(x*scale)
void launch_scale_convert(int32_t n, const double* x, double scale, float* y, hipStream_t s);
// psi += scale*(double)d
void launch_axpy_f32(int32_t n, double* psi, const float* d, double scale, hipStream_t s);
// y -= alpha*x with the state's alpha (PBiCG transpose residual)
template<class T>
void launch_axpy_alpha(int32_t n, T* y, const T* x, const DevState* st, hipStream_t s);
// r = source - Apsi, sum |r| into the second partial (GAMGSolverSolve.C:120-128)
void launch_sub_sum_abs(int32_t n, const double* Apsi, const double* source, double* r,
                        double* partials, const DevState* st, hipStream_t s);
// psi = src/diag
void launch_divide(int32_t n, const double* src, const double* diag, double* psi,
                   hipStream_t s);

} // namespace fh
