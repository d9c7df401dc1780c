// GAMG on the device: agglomeration hierarchy, Galerkin coarse matrices,
// V-cycle (GAMGSolverSolve.C:31-552) and the GAMG preconditioner.
#pragma once

#include "fh_matrix.h"

#include <vector>

namespace fh
{
// coefficients changed: coarse matrices must be re-agglomerated
void gamg_invalidate(fh_matrix* M);
// free the matrix's coarse levels (the mesh keeps its agglomeration)
void gamg_release(fh_matrix* M);
// build (or reuse) the hierarchy for these controls and bring the coarse
// coefficients up to date
void gamg_prepare(fh_matrix* M, const fh_controls* c);
// the faceAreaPair weights of a mesh's internal faces (drops its cached
// agglomeration)
void gamg_set_face_weights(fh_mesh* mesh, std::vector<double>&& w);
// GAMG preconditioner: w = nVcycles V-cycles on A w = r from w = 0
template<class T>
void gamg_apply(fh_matrix* M, const T* r, T* w, DevState* st);
// GAMG solver (FP64, or FP32 V-cycles inside FP64 residual updates)
void gamg_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                fh_perf* perf, bool mixed);
// FH_ERR_SCHEDULE when a sweep wait of a coarse level expired
void gamg_check_sweep_errors(fh_matrix* M);
// pairGAMGAgglomeration::forward_ access (process-global)
bool gamg_forward();
void gamg_set_forward(bool f);
} // namespace fh
