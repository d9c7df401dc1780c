// GAMG on the device (GAMGSolver*.C, GAMGAgglomeration*.C).
#include "fh_gamg.h"

namespace fh
{

// pairGAMGAgglomeration::forward_ (pairGAMGAgglomeration.C:33): a
// process-global direction toggle flipped after every pair agglomeration
bool g_forward = true;

void gamg_invalidate(fh_matrix*) {}

void gamg_solve(fh_matrix*, const fh_controls*, double*, const double*, fh_perf*, bool)
{
    fail(FH_ERR_UNSUPPORTED, "GAMG is not available in this build yet");
}

void gamg_pcg(fh_matrix*, const fh_controls*, double*, double*, double*, double*)
{
    fail(FH_ERR_UNSUPPORTED, "the GAMG preconditioner is not available in this build yet");
}

void pbicg_solve(fh_matrix*, const fh_controls*, double*, const double*, fh_perf*)
{
    fail(FH_ERR_UNSUPPORTED, "PBiCG is not available in this build yet");
}

} // namespace fh

extern "C" {
int fh_gamg_direction(void) { return fh::g_forward ? 1 : 0; }
void fh_gamg_reset_direction(void) { fh::g_forward = true; }
}
