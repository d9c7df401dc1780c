// Coarse correction of the tile-local preconditioner (fh_coarse.hip): the
// coarse matrix over 16^3 tiles, its dense FP32 inverse and its apply.
#pragma once

#include "fh_kernels.h"
#include "fh_solver_kernels.h"

namespace fh
{

// per tile t: ctile[4t] = diagonal sum + twice the inner couplings,
// ctile[4t+1..3] = couplings across the tile's lower x, y, z faces
void launch_coarse_tiles(int32_t nx, int32_t ny, int32_t nz, int32_t nTI, int32_t nTJ,
                         int32_t nT, const double* diag, const double* const dir[3],
                         double* ctile, hipStream_t s);
// dense (column-major) s*Ac, s the sign of the first diagonal entry
// (positive definite for a definite matrix of either sign)
void launch_coarse_dense(int32_t nT, int32_t nTI, int32_t nTJ, const double* ctile,
                         double* dense, hipStream_t s);
// full FP32 inverse of Ac from the lower triangle of the factored inverse of
// s*Ac, rows of ld floats
void launch_coarse_inv32(int32_t nT, int32_t ld, const double* dense, const double* ctile,
                         float* inv, hipStream_t s);
// the FP32 inverse of Ac through its block-tridiagonal structure over the
// planes of P = nTI*nTJ tiles (P <= 256, at least two planes; returns false
// and launches nothing otherwise): U holds the K = nT/P Schur complement
// inverses (K*P*P doubles), Z the dense FP64 inverse (nT*nT), work
// coarse_block_work_doubles(nT, P) doubles; *bad (zeroed by the caller)
// becomes 1 if the matrix is not definite
bool launch_coarse_block_inverse(int32_t nT, int32_t nTI, int32_t nTJ, int32_t ld,
                                 const double* ctile, double* U, double* Z, double* work,
                                 float* inv, int32_t* bad, hipStream_t s);
int64_t coarse_block_work_doubles(int32_t nT, int32_t P);
// z = inv rc, and sum z*rc plus the nF fine partials of sum w*r, finalized
// by the last block with F (FIN_PCG_WARA); returns the block count
int32_t launch_coarse_apply(int32_t nT, int32_t ld, const float* inv, const double* rc,
                            float* z, const double* fine, int32_t nF, double* partials,
                            const FusedFinalize& F, const DevState* st, hipStream_t s);
// w += z of the cell's 16^3 tile (the coarse correction added to the
// tile-local apply where the product does not add it)
void launch_add_tile_correction(int32_t nx, int32_t ny, int32_t nz, const float* z, float* w,
                                const DevState* st, hipStream_t s);
// p = (w + z) + beta p with w on rows padded to nxp cells (z may be null)
void launch_pcg_p_padded(int32_t nx, int32_t ny, int32_t nz, int32_t nxp, const float* wp,
                         const float* z, float* p, const DevState* st, hipStream_t s);

} // namespace fh
