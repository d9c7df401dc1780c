#include "hipLduSolver.H"
#include "addToRunTimeSelectionTable.H"
#include "fvMesh.H"

#include <map>
#include <set>

namespace Foam
{
    defineTypeNameAndDebug(hipPCG, 0);
    lduMatrix::solver::addsymMatrixConstructorToTable<hipPCG>
        addhipPCGSymMatrixConstructorToTable_;

    defineTypeNameAndDebug(hipPBiCGStab, 0);
    lduMatrix::solver::addsymMatrixConstructorToTable<hipPBiCGStab>
        addhipPBiCGStabSymMatrixConstructorToTable_;
    lduMatrix::solver::addasymMatrixConstructorToTable<hipPBiCGStab>
        addhipPBiCGStabAsymMatrixConstructorToTable_;

    defineTypeNameAndDebug(hipPBiCG, 0);
    lduMatrix::solver::addasymMatrixConstructorToTable<hipPBiCG>
        addhipPBiCGAsymMatrixConstructorToTable_;

    defineTypeNameAndDebug(hipSmoothSolver, 0);
    lduMatrix::solver::addsymMatrixConstructorToTable<hipSmoothSolver>
        addhipSmoothSolverSymMatrixConstructorToTable_;
    lduMatrix::solver::addasymMatrixConstructorToTable<hipSmoothSolver>
        addhipSmoothSolverAsymMatrixConstructorToTable_;

    defineTypeNameAndDebug(hipGAMG, 0);
    lduMatrix::solver::addsymMatrixConstructorToTable<hipGAMG>
        addhipGAMGSymMatrixConstructorToTable_;
    lduMatrix::solver::addasymMatrixConstructorToTable<hipGAMG>
        addhipGAMGAsymMatrixConstructorToTable_;
}


Foam::hipLduSolver::hipLduSolver
(
    const word& fieldName,
    const lduMatrix& matrix,
    const Field<Field<scalar>>& interfaceBouCoeffs,
    const Field<Field<scalar>>& interfaceIntCoeffs,
    const lduInterfaceFieldPtrsList& interfaces,
    const dictionary& solverControls,
    const int32_t solverId
)
:
    lduMatrix::solver
    (
        fieldName,
        matrix,
        interfaceBouCoeffs,
        interfaceIntCoeffs,
        interfaces,
        solverControls
    ),
    solverId_(solverId)
{}


void Foam::hipLduSolver::check(int status, const char* what)
{
    if (status != FH_OK)
    {
        FatalErrorInFunction
            << what << " failed (status " << status << "): "
            << fh_last_error() << exit(FatalError);
    }
}


fh_mesh* Foam::hipLduSolver::deviceMesh(const lduAddressing& addr)
{
    // keyed by the addressing object and its size; a topology change
    // produces a new lduAddressing
    static std::map<std::pair<const lduAddressing*, label>, fh_mesh*> meshes;
    const auto key = std::make_pair(&addr, addr.size());
    auto it = meshes.find(key);
    if (it != meshes.end()) return it->second;
    fh_mesh* m = nullptr;
    check
    (
        fh_mesh_create
        (
            addr.size(),
            addr.lowerAddr().size(),
            addr.lowerAddr().begin(),
            addr.upperAddr().begin(),
            &m
        ),
        "fh_mesh_create"
    );
    meshes[key] = m;
    return m;
}


fh_matrix* Foam::hipLduSolver::deviceMatrix(fh_mesh* mesh)
{
    static std::map<fh_mesh*, fh_matrix*> matrices;
    auto it = matrices.find(mesh);
    if (it != matrices.end()) return it->second;
    fh_matrix* m = nullptr;
    check(fh_matrix_create(mesh, &m), "fh_matrix_create");
    matrices[mesh] = m;
    return m;
}


void Foam::hipLduSolver::readControls(const dictionary& dict, fh_controls& c)
{
    fh_controls_default(&c);

    // lduMatrixSolver.C:158-164
    c.maxIter = dict.lookupOrDefault<label>("maxIter", 1000);
    c.minIter = dict.lookupOrDefault<label>("minIter", 0);
    c.tolerance = dict.lookupOrDefault<scalar>("tolerance", 1e-6);
    c.relTol = dict.lookupOrDefault<scalar>("relTol", 0);

    // preconditioner: word or sub-dictionary (lduMatrixPreconditioner.C:40-60)
    if (dict.found("preconditioner"))
    {
        const word name(lduMatrix::preconditioner::getName(dict));
        check(fh_preconditioner_lookup(name.c_str(), &c.preconditioner), "preconditioner");
    }
    if (dict.found("smoother"))
    {
        const word name(dict.lookup("smoother"));
        check(fh_smoother_lookup(name.c_str(), &c.smoother), "smoother");
    }

    // smoothSolver.C:70-74
    c.nSweeps = dict.lookupOrDefault<label>("nSweeps", 1);

    // GAMGSolver.C:351-388, GAMGAgglomeration.C:248-257, pairGAMGAgglomeration.C:44-52
    c.nPreSweeps = dict.lookupOrDefault<label>("nPreSweeps", 0);
    c.preSweepsLevelMultiplier = dict.lookupOrDefault<label>("preSweepsLevelMultiplier", 1);
    c.maxPreSweeps = dict.lookupOrDefault<label>("maxPreSweeps", 4);
    c.nPostSweeps = dict.lookupOrDefault<label>("nPostSweeps", 2);
    c.postSweepsLevelMultiplier = dict.lookupOrDefault<label>("postSweepsLevelMultiplier", 1);
    c.maxPostSweeps = dict.lookupOrDefault<label>("maxPostSweeps", 4);
    c.nFinestSweeps = dict.lookupOrDefault<label>("nFinestSweeps", 2);
    if (dict.found("scaleCorrection"))
    {
        c.scaleCorrection = readBool(dict.lookup("scaleCorrection"));
    }
    c.directSolveCoarsest = dict.lookupOrDefault<bool>("directSolveCoarsest", false);
    c.nCellsInCoarsestLevel = dict.lookupOrDefaultBackwardsCompatible<label>
    (
        {"minCellsPerProcessor", "nCellsInCoarsestLevel"},
        10
    );
    c.mergeLevels = dict.lookupOrDefault<label>("mergeLevels", 1);
    c.nVcycles = dict.lookupOrDefault<label>("nVcycles", 2);

    // GAMGAgglomeration::New (GAMGAgglomeration.C:291-340): faceAreaPair
    // by default, its face areas handed to the device mesh in solve()
    const word agg(dict.lookupOrDefault<word>("agglomerator", "faceAreaPair"));
    if (agg == "faceAreaPair")
    {
        c.agglomerator = FH_AGGLOM_FACE_AREA_PAIR;
    }
    else if (agg == "algebraicPair")
    {
        c.agglomerator = FH_AGGLOM_ALGEBRAIC_PAIR;
    }
    else
    {
        FatalIOErrorInFunction(dict)
            << "Unsupported agglomerator " << agg
            << "; valid: algebraicPair faceAreaPair" << exit(FatalIOError);
    }

    const word precision(dict.lookupOrDefault<word>("precision", "double"));
    if (precision == "mixed")
    {
        c.precision = FH_PRECISION_MIXED;
    }
    else if (precision != "double")
    {
        FatalIOErrorInFunction(dict)
            << "Unknown precision " << precision << "; valid: double mixed"
            << exit(FatalIOError);
    }
}


Foam::solverPerformance Foam::hipLduSolver::solve
(
    scalarField& psi,
    const scalarField& source,
    const direction
) const
{
    // coupled interfaces (processor/cyclic patches) are outside this
    // single-GPU path
    forAll(interfaces_, patchi)
    {
        if (interfaces_.set(patchi))
        {
            FatalErrorInFunction
                << type() << " does not handle coupled interfaces (patch "
                << patchi << "); use the host solver for this field"
                << exit(FatalError);
        }
    }

    fh_controls c;
    readControls(controlDict_, c);
    c.solver = solverId_;

    fh_mesh* mesh = deviceMesh(matrix_.lduAddr());
    fh_matrix* m = deviceMatrix(mesh);

    // faceAreaPair weights from fvMesh::Sf() and magSf()
    // (faceAreaPairGAMGAgglomeration.C:55-80), once per device mesh
    static std::set<fh_mesh*> withFaceAreas;
    const bool gamg =
        solverId_ == FH_SOLVER_GAMG || c.preconditioner == FH_PRECON_GAMG;
    if
    (
        gamg
     && c.agglomerator == FH_AGGLOM_FACE_AREA_PAIR
     && !withFaceAreas.count(mesh)
    )
    {
        const fvMesh& fvm = refCast<const fvMesh>(matrix_.mesh());
        const vectorField& Sf = fvm.Sf().primitiveField();
        const scalarField& magSf = fvm.magSf().primitiveField();
        check
        (
            fh_mesh_set_face_areas
            (
                mesh,
                reinterpret_cast<const scalar*>(Sf.cdata()),
                magSf.cdata()
            ),
            "fh_mesh_set_face_areas"
        );
        withFaceAreas.insert(mesh);
    }

    const scalar* upper = matrix_.hasUpper() ? matrix_.upper().begin() : nullptr;
    const scalar* lower =
        matrix_.asymmetric() ? matrix_.lower().begin() : nullptr;
    check(fh_matrix_set(m, matrix_.diag().begin(), upper, lower), "fh_matrix_set");

    fh_perf perf;
    check(fh_solve(m, &c, psi.begin(), source.begin(), &perf), "fh_solve");

    // "<preconditioner><solver>" naming of PCG.C:73-77, with the hip prefix
    solverPerformance solverPerf(word(perf.solverName), fieldName_);
    solverPerf.initialResidual() = perf.initialResidual;
    solverPerf.finalResidual() = perf.finalResidual;
    solverPerf.nIterations() = perf.nIterations;
    solverPerf.checkConvergence(tolerance_, relTol_);
    if (perf.singular)
    {
        solverPerf.checkSingularity(0);
    }
    return solverPerf;
}
