/*---------------------------------------------------------------------------*\
    hipDIC / hipDILU: preconditioners that apply DIC / DILU on the device for
    a host-side Krylov solver (DICPreconditioner.C:87-123,
    DILUPreconditioner.C:88-185).  Each application copies rA to the device
    and wA back, so it is correct but transfer-bound; pair hipPCG with
    "preconditioner DIC" instead for a device-resident solve.
\*---------------------------------------------------------------------------*/
#ifndef hipLduPreconditioner_H
#define hipLduPreconditioner_H

#include "lduMatrix.H"

extern "C" {
#include "foamhip.h"
}

namespace Foam
{

class hipLduPreconditioner
:
    public lduMatrix::preconditioner
{
    const int32_t preconId_;
    fh_matrix* matrix_;

public:

    hipLduPreconditioner
    (
        const lduMatrix::solver& sol,
        const dictionary& dict,
        const int32_t preconId
    );

    virtual void precondition
    (
        scalarField& wA,
        const scalarField& rA,
        const direction cmpt = 0
    ) const;

    virtual void preconditionT
    (
        scalarField& wT,
        const scalarField& rT,
        const direction cmpt = 0
    ) const;
};


#define declareHipPreconditioner(Type, Id)                                    \
class Type                                                                    \
:                                                                             \
    public hipLduPreconditioner                                               \
{                                                                             \
public:                                                                       \
    TypeName(#Type);                                                          \
    Type(const lduMatrix::solver& sol, const dictionary& dict)                \
    :                                                                         \
        hipLduPreconditioner(sol, dict, Id)                                   \
    {}                                                                        \
};

declareHipPreconditioner(hipDIC, FH_PRECON_DIC)
declareHipPreconditioner(hipDILU, FH_PRECON_DILU)

} // End namespace Foam

#endif
