#include "hipLduPreconditioner.H"
#include "hipLduSolver.H"

namespace Foam
{
    defineTypeNameAndDebug(hipDIC, 0);
    lduMatrix::preconditioner::addsymMatrixConstructorToTable<hipDIC>
        addhipDICPreconditionerSymMatrixConstructorToTable_;

    defineTypeNameAndDebug(hipDILU, 0);
    lduMatrix::preconditioner::addasymMatrixConstructorToTable<hipDILU>
        addhipDILUPreconditionerAsymMatrixConstructorToTable_;
}


Foam::hipLduPreconditioner::hipLduPreconditioner
(
    const lduMatrix::solver& sol,
    const dictionary&,
    const int32_t preconId
)
:
    lduMatrix::preconditioner(sol),
    preconId_(preconId),
    matrix_(nullptr)
{
    const lduMatrix& m = sol.matrix();
    fh_mesh* mesh = hipLduSolver::deviceMesh(m.lduAddr());
    fh_matrix* dm = hipLduSolver::deviceMatrix(mesh);
    const scalar* lower = m.asymmetric() ? m.lower().begin() : nullptr;
    hipLduSolver::check
    (
        fh_matrix_set(dm, m.diag().begin(), m.upper().begin(), lower),
        "fh_matrix_set"
    );
    matrix_ = dm;
}


void Foam::hipLduPreconditioner::precondition
(
    scalarField& wA,
    const scalarField& rA,
    const direction
) const
{
    hipLduSolver::check
    (
        fh_precondition(matrix_, preconId_, rA.begin(), wA.begin()),
        "fh_precondition"
    );
}


void Foam::hipLduPreconditioner::preconditionT
(
    scalarField& wT,
    const scalarField& rT,
    const direction
) const
{
    hipLduSolver::check
    (
        fh_precondition_t(matrix_, preconId_, rT.begin(), wT.begin()),
        "fh_precondition_t"
    );
}
