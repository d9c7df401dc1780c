// lduMatrix H, H1, faceH and sumMagOffDiag on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "fh_kernels.h"
#include "fh_assembly.h"

namespace fh
{

// y = H(x) over the library rows; x and y hold nC components per cell,
// component fastest (1 <= nC <= 9)
void launch_h(const RowsView& R, const double* val, int32_t nC, const double* x, double* y,
              hipStream_t s);
// y = H1 over the library rows
void launch_h1(const RowsView& R, const double* val, double* y, hipStream_t s);
// y += sumMagOffDiag over the library rows
void launch_sum_mag_off_diag(const RowsView& R, const double* val, double* y, hipStream_t s);
// y[f] = upper[f]*x[u[f]] - lower[f]*x[l[f]] per face (natural numbering)
void launch_face_h(int32_t nFaces, const int32_t* l, const int32_t* u, const double* upper,
                   const double* lower, int32_t nC, const double* x, double* y, hipStream_t s);
// fvMatrix<Type>::relax(alpha) with non-coupled patches: diag (library
// numbering) replaced, source (natural, nC components) updated; diagOut
// (natural, may be null) receives the relaxed diagonal
void launch_relax(const RowsView& R, const int32_t* perm, const double* val,
                  const BoundaryView& B, const double* ic, int32_t nC, double alpha,
                  const double* psi, double* diag, double* source, double* diagOut,
                  hipStream_t s);
// fvMatrix<Type>::flux() with non-coupled patches: y (nFaces x nC) and
// by (nB x nC); corr, bCorr: the face flux correction or null
void launch_flux(int32_t nFaces, const int32_t* l, const int32_t* u, const double* upper,
                 const double* lower, const BoundaryView& B, int32_t nC, const double* x,
                 const double* ic, const double* bc, const double* corr, const double* bCorr,
                 double* y, double* by, hipStream_t s);
// component-wise permutations between the natural and library numbering
void launch_permute_cmpt_in(int32_t n, int32_t nC, const int32_t* perm, const double* xNat,
                            double* xLib, hipStream_t s);
void launch_permute_cmpt_out(int32_t n, int32_t nC, const int32_t* perm, const double* xLib,
                             double* xNat, hipStream_t s);

} // namespace fh
