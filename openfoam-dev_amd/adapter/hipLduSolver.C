#include "hipLduSolver.H"
#include "addToRunTimeSelectionTable.H"
#include "fvMesh.H"
#include "surfaceFields.H"
#include "cyclicLduInterface.H"

#include <map>

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


namespace
{
    // FNV-1a over a block of memory: identifies the contents of addressing
    // and face-area arrays, so a new mesh at a reused address, or moved
    // points, is not mistaken for the cached one
    uint64_t contentHash(const void* data, size_t bytes, uint64_t h = 1469598103934665603ull)
    {
        const unsigned char* p = static_cast<const unsigned char*>(data);
        for (size_t i = 0; i < bytes; i++)
        {
            h = (h ^ p[i])*1099511628211ull;
        }
        return h;
    }

    struct cachedMesh
    {
        Foam::label nCells;
        Foam::label nFaces;
        uint64_t hash;
        fh_mesh* mesh;
    };

    std::map<fh_mesh*, fh_matrix*>& matrixCache()
    {
        static std::map<fh_mesh*, fh_matrix*> matrices;
        return matrices;
    }
}


fh_mesh* Foam::hipLduSolver::deviceMesh(const lduAddressing& addr)
{
    return deviceMesh(addr, labelList(), labelList(), labelList());
}


fh_mesh* Foam::hipLduSolver::deviceMesh
(
    const lduAddressing& addr,
    const labelList& interfaceStart,
    const labelList& faceCells,
    const labelList& nbrFaceCells
)
{
    // keyed by the addressing object; a hit is checked against the cell
    // and face counts and the contents of lowerAddr/upperAddr, and a stale
    // entry (topology change, or another addressing allocated at the same
    // address) is destroyed with its matrix and rebuilt
    static std::map<const lduAddressing*, cachedMesh> meshes;
    const labelUList& l = addr.lowerAddr();
    const labelUList& u = addr.upperAddr();
    uint64_t h = contentHash
    (
        u.begin(),
        u.size()*sizeof(label),
        contentHash(l.begin(), l.size()*sizeof(label))
    );
    h = contentHash(faceCells.cdata(), faceCells.size()*sizeof(label), h);
    h = contentHash(nbrFaceCells.cdata(), nbrFaceCells.size()*sizeof(label), h);
    auto it = meshes.find(&addr);
    if (it != meshes.end())
    {
        const cachedMesh& cm = it->second;
        if (cm.nCells == addr.size() && cm.nFaces == l.size() && cm.hash == h)
        {
            return cm.mesh;
        }
        auto mit = matrixCache().find(cm.mesh);
        if (mit != matrixCache().end())
        {
            fh_matrix_destroy(mit->second);
            matrixCache().erase(mit);
        }
        fh_mesh_destroy(cm.mesh);
        meshes.erase(it);
    }
    fh_mesh* m = nullptr;
    check
    (
        fh_mesh_create_coupled
        (
            addr.size(),
            l.size(),
            l.begin(),
            u.begin(),
            interfaceStart.size() ? interfaceStart.size() - 1 : 0,
            interfaceStart.cdata(),
            faceCells.cdata(),
            nbrFaceCells.cdata(),
            &m
        ),
        "fh_mesh_create_coupled"
    );
    meshes[&addr] = cachedMesh{addr.size(), l.size(), h, m};
    return m;
}


fh_matrix* Foam::hipLduSolver::deviceMatrix(fh_mesh* mesh)
{
    std::map<fh_mesh*, fh_matrix*>& matrices = matrixCache();
    auto it = matrices.find(mesh);
    if (it != matrices.end()) return it->second;
    fh_matrix* m = nullptr;
    check(fh_matrix_create(mesh, &m), "fh_matrix_create");
    matrices[mesh] = m;
    return m;
}


void Foam::hipLduSolver::readControls
(
    const dictionary& dict,
    fh_controls& c,
    fh_controls& gamgPrecond
)
{
    readControls(dict, c);

    // a GAMG preconditioner is a GAMG solver built from the preconditioner
    // sub-dictionary alone (lduMatrixPreconditioner.C:72-82,
    // GAMGPreconditioner.C:44-63); the word form hands it dictionary::null,
    // which has no smoother entry
    if (c.preconditioner == FH_PRECON_GAMG)
    {
        const entry& e = dict.lookupEntry("preconditioner", false, false);
        if (!e.isDict())
        {
            FatalIOErrorInFunction(dict)
                << "preconditioner GAMG reads its controls from a preconditioner"
                << " sub-dictionary: keyword smoother is undefined"
                << exit(FatalIOError);
        }
        readControls(e.dict(), gamgPrecond);
        gamgPrecond.solver = FH_SOLVER_GAMG;
        gamgPrecond.preconditioner = FH_PRECON_NONE;
        c.preconditionerControls = &gamgPrecond;
    }
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
    // coupled interfaces: cyclic patches without a transformation of the
    // field (cyclicFvPatchField::updateInterfaceMatrix with pnf from the
    // neighbour patch's faceCells, cyclicFvPatchField.C:150-174) go to the
    // device; processor patches (another rank's cells) and transforming
    // cyclics are outside this single-GPU path
    labelList ifStart(1, 0), ifCells, ifNbr;
    DynamicList<scalar> ifBou, ifInt;
    forAll(interfaces_, patchi)
    {
        if (!interfaces_.set(patchi)) continue;
        const lduInterface& intf = interfaces_[patchi].interface();
        if
        (
            !isA<cyclicLduInterface>(intf)
         || refCast<const cyclicLduInterface>(intf).transform().transforms()
        )
        {
            FatalErrorInFunction
                << type() << " handles non-transforming cyclic interfaces only (patch "
                << patchi << " is " << intf.type() << "); use the host solver for this field"
                << exit(FatalError);
        }
        const label nbr = refCast<const cyclicLduInterface>(intf).nbrPatchIndex();
        const labelUList& fc = intf.faceCells();
        const labelUList& nfc = interfaces_[nbr].interface().faceCells();
        forAll(fc, i)
        {
            ifCells.append(fc[i]);
            ifNbr.append(nfc[i]);
            ifBou.append(interfaceBouCoeffs_[patchi][i]);
            ifInt.append(interfaceIntCoeffs_[patchi][i]);
        }
        ifStart.append(ifCells.size());
    }
    if (ifStart.size() == 1) ifStart.clear();

    fh_controls c, gamgPrecond;
    readControls(controlDict_, c, gamgPrecond);
    c.solver = solverId_;

    fh_mesh* mesh = deviceMesh(matrix_.lduAddr(), ifStart, ifCells, ifNbr);
    fh_matrix* m = deviceMatrix(mesh);

    // faceAreaPair weights from fvMesh::Sf() and magSf()
    // (faceAreaPairGAMGAgglomeration.C:55-80); handed over again when the
    // face areas change (mesh motion), which drops the device mesh's cached
    // agglomeration as movePoints drops the reference's GeometricMeshObject
    static std::map<fh_mesh*, uint64_t> faceAreaHash;
    const fh_controls& gc =
        c.preconditionerControls ? *c.preconditionerControls : c;
    const bool gamg =
        solverId_ == FH_SOLVER_GAMG || c.preconditioner == FH_PRECON_GAMG;
    if (gamg && gc.agglomerator == FH_AGGLOM_FACE_AREA_PAIR)
    {
        const fvMesh& fvm = refCast<const fvMesh>(matrix_.mesh());
        const vectorField& Sf = fvm.Sf().primitiveField();
        const scalarField& magSf = fvm.magSf().primitiveField();
        const uint64_t h = contentHash
        (
            magSf.cdata(),
            magSf.size()*sizeof(scalar),
            contentHash(Sf.cdata(), Sf.size()*sizeof(vector))
        );
        auto it = faceAreaHash.find(mesh);
        if (it == faceAreaHash.end() || it->second != h)
        {
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
            faceAreaHash[mesh] = h;
        }
    }

    const scalar* upper = matrix_.hasUpper() ? matrix_.upper().begin() : nullptr;
    const scalar* lower =
        matrix_.asymmetric() ? matrix_.lower().begin() : nullptr;
    check(fh_matrix_set(m, matrix_.diag().begin(), upper, lower), "fh_matrix_set");
    if (ifBou.size())
    {
        check
        (
            fh_matrix_set_interfaces(m, ifBou.cdata(), ifInt.cdata()),
            "fh_matrix_set_interfaces"
        );
    }

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
