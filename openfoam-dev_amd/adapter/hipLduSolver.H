/*---------------------------------------------------------------------------*\
    hipLduSolver: run-time-selectable lduMatrix solvers that solve on an
    MI355X through the foamhip C-ABI (include/foamhip.h).

    Registered names (fvSolution "solver" entry):
        hipPCG        symMatrix table   (PCG.C:34-35)
        hipPBiCGStab  sym + asym tables (PBiCGStab.C:34-38)
        hipPBiCG      asymMatrix table  (PBiCG.C:34-35)
        hipSmoothSolver, hipGAMG  sym + asym tables
    Extra optional entry:  precision double|mixed;  (default double)

    The solver keeps lduMatrix::solver's constructor signature
    (lduMatrix.H:187-195) and solve() contract (lduMatrix.H:250-255): psi is
    the initial guess on entry and the solution on return.  Device addressing
    is built once per lduAddressing and reused; coefficients are uploaded on
    every solve because fvMatrix::solveSegregated adds the boundary
    coefficients to the diagonal only for the duration of the call
    (fvScalarMatrix.C:155-177).
\*---------------------------------------------------------------------------*/
#ifndef hipLduSolver_H
#define hipLduSolver_H

#include "lduMatrix.H"

extern "C" {
#include "foamhip.h"
}

namespace Foam
{

class hipLduSolver
:
    public lduMatrix::solver
{
protected:

    //- foamhip solver id of this run-time type
    const int32_t solverId_;

public:

    hipLduSolver
    (
        const word& fieldName,
        const lduMatrix& matrix,
        const Field<Field<scalar>>& interfaceBouCoeffs,
        const Field<Field<scalar>>& interfaceIntCoeffs,
        const lduInterfaceFieldPtrsList& interfaces,
        const dictionary& solverControls,
        const int32_t solverId
    );

    virtual ~hipLduSolver() {}

    //- Solve on the device; same contract as lduMatrix::solver::solve
    virtual solverPerformance solve
    (
        scalarField& psi,
        const scalarField& source,
        const direction cmpt = 0
    ) const;

    //- Device mesh of an lduAddressing, built on first use and cached
    static fh_mesh* deviceMesh(const lduAddressing& addr);
    //- The device mesh with the matrix's coupled (cyclic) interfaces:
    //  interfaceStart, faceCells and nbrFaceCells concatenated in interface
    //  order (empty: no interfaces)
    static fh_mesh* deviceMesh
    (
        const lduAddressing& addr,
        const labelList& interfaceStart,
        const labelList& faceCells,
        const labelList& nbrFaceCells
    );

    //- Device matrix of a mesh (coefficients replaced on every solve)
    static fh_matrix* deviceMatrix(fh_mesh* mesh);

    //- fvSolution dictionary -> fh_controls
    static void readControls(const dictionary& dict, fh_controls& c);

    //- The same, and for a GAMG preconditioner its sub-dictionary's
    //  controls into gamgPrecond (pointed to by c.preconditionerControls)
    static void readControls
    (
        const dictionary& dict,
        fh_controls& c,
        fh_controls& gamgPrecond
    );

    //- Turn a foamhip status into FatalError
    static void check(int status, const char* what);
};


#define declareHipSolver(Type, Id)                                            \
class Type                                                                    \
:                                                                             \
    public hipLduSolver                                                       \
{                                                                             \
public:                                                                       \
    TypeName(#Type);                                                          \
    Type                                                                      \
    (                                                                         \
        const word& fieldName,                                                \
        const lduMatrix& matrix,                                              \
        const Field<Field<scalar>>& interfaceBouCoeffs,                  \
        const Field<Field<scalar>>& interfaceIntCoeffs,                  \
        const lduInterfaceFieldPtrsList& interfaces,                          \
        const dictionary& solverControls                                      \
    )                                                                         \
    :                                                                         \
        hipLduSolver                                                          \
        (                                                                     \
            fieldName, matrix, interfaceBouCoeffs, interfaceIntCoeffs,        \
            interfaces, solverControls, Id                                    \
        )                                                                     \
    {}                                                                        \
};

declareHipSolver(hipPCG, FH_SOLVER_PCG)
declareHipSolver(hipPBiCGStab, FH_SOLVER_PBICGSTAB)
declareHipSolver(hipPBiCG, FH_SOLVER_PBICG)
declareHipSolver(hipSmoothSolver, FH_SOLVER_SMOOTH)
declareHipSolver(hipGAMG, FH_SOLVER_GAMG)

} // End namespace Foam

#endif
