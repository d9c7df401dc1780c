// GAMG on the device: agglomeration hierarchy, Galerkin coarse matrices,
// V-cycle (GAMGSolverSolve.C:31-552) and the GAMG preconditioner.
#pragma once

#include "fh_matrix.h"

namespace fh
{
void gamg_invalidate(fh_matrix* M);
void gamg_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                fh_perf* perf, bool mixed);
void gamg_pcg(fh_matrix* M, const fh_controls* c, double* psi, double* rA, double* pA,
              double* wA);
void pbicg_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                 fh_perf* perf);
} // namespace fh
