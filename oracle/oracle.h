/*
 * oracle.h -- CPU restatement of OpenFOAM-dev's lduMatrix linear-solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it; the product path never
 * links or calls it.
 *
 * Every routine follows the loop order of the reference function it names
 * (citations are relative to the OpenFOAM-dev source tree), so that results
 * are bitwise reproducible in plain FP64 with no FMA contraction
 * (compile with -O3 -ffp-contract=off, no -ffast-math).
 *
 * Scope: serial (single-rank) solves; coupled interfaces (non-transforming
 * cyclic patches, or_interfaces) where a matrix is given them.
 */
#ifndef FOAMHIP_ORACLE_H
#define FOAMHIP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LDU addressing (lduAddressing.H:112-207, lduAddressing.C:32-170) ---- */
typedef struct
{
    int32_t nCells;
    int32_t nFaces;
    const int32_t* lower;   /* owner     (lowerAddr) [nFaces] */
    const int32_t* upper;   /* neighbour (upperAddr) [nFaces] */
    int32_t* losort;        /* [nFaces]   */
    int32_t* ownerStart;    /* [nCells+1] */
    int32_t* losortStart;   /* [nCells+1] */
} or_addr;

int  or_addr_init(or_addr* a, int32_t nCells, int32_t nFaces,
                  const int32_t* lower, const int32_t* upper);
void or_addr_free(or_addr* a);

/* Coupled interfaces of a matrix (lduInterfaceFieldPtrsList with their
   interfaceBouCoeffs / interfaceIntCoeffs): interface i owns faces
   start[i] .. start[i+1]-1 of the concatenated arrays; face e couples
   faceCells[e] to nbrCells[e] (cyclicFvPatchField::updateInterfaceMatrix,
   cyclicFvPatchField.C:150-174: result[faceCells] -= coeffs*psi[nbr
   faceCells], no transformation for scalars). */
typedef struct
{
    int32_t nInterfaces;
    const int32_t* start;       /* [nInterfaces+1] */
    const int32_t* faceCells;   /* [start[nInterfaces]] */
    const int32_t* nbrCells;
    const double* bouCoeffs;    /* interfaceBouCoeffs */
    const double* intCoeffs;    /* interfaceIntCoeffs (Tmul) */
} or_interfaces;

/* Matrix view: lower == upper (same pointer) for a symmetric matrix,
   as lduMatrix::lower() const returns upper when lowerPtr_ is unset
   (lduMatrix.C:168-312); interfaces NULL: no coupled patches. */
typedef struct
{
    const or_addr* addr;
    const double* diag;
    const double* upper;
    const double* lower;
    const or_interfaces* interfaces;
} or_matrix;

/* ---- reductions (FieldReductionFunctions.C) ---- */
double or_sum(int32_t n, const double* f);
double or_sumMag(int32_t n, const double* f);
double or_sumSqr(int32_t n, const double* f);
double or_sumProd(int32_t n, const double* a, const double* b);
double or_average(int32_t n, const double* f);
/* 0: serial (the reference), 1: blocks of 256 then the block sums (a
   different order of the same additions, for measuring sensitivity) */
void or_set_sum_order(int order);

/* ---- matrix kernels (lduMatrixATmul.C, lduMatrixOperations.C) ---- */
void or_amul(const or_matrix* m, const double* psi, double* Apsi);
void or_tmul(const or_matrix* m, const double* psi, double* Tpsi);
void or_sumA(const or_matrix* m, double* sumA);
void or_residual(const or_matrix* m, const double* psi, const double* source,
                 double* rA);
double or_normFactor(const or_matrix* m, const double* psi,
                     const double* source, const double* Apsi, double* tmp);
void or_negSumDiag(const or_addr* a, const double* lower, const double* upper,
                   double* diag);

/* ---- the off-diagonal face loops of the pressure-velocity coupling ----
   Fields of a Type with nCmpt components are stored component fastest
   (the memory layout of Field<Type>); m->upper == NULL is a diagonal-only
   matrix. */
/* lduMatrixTemplates.C:33-64 */
void or_H(const or_matrix* m, int32_t nCmpt, const double* psi, double* Hpsi);
/* lduMatrixATmul.C:298-324 */
void or_H1(const or_matrix* m, double* H1);
/* lduMatrixTemplates.C:77-111; returns -1 (the reference's fatal error)
   when the matrix has no off-diagonal coefficients, else 0 */
int or_faceH(const or_matrix* m, int32_t nCmpt, const double* psi, double* faceHpsi);
/* lduMatrixOperations.C:77-97: sumOff is added to */
void or_sumMagOffDiag(const or_matrix* m, double* sumOff);
/* fvMatrix.C:953-1028, fvMatrix<Type>::flux() with every patch
   non-coupled: flux (nFaces x nCmpt) = faceH (+ corr), bFlux (nB x nCmpt)
   = ic * psi(faceCell) - bc (+ bCorr); corr, bCorr may be NULL; returns
   -1 for a diagonal-only matrix */
int or_flux(const or_matrix* m, int32_t nCmpt, const double* psi, int32_t nB,
            const int32_t* bFaceCells, const double* ic, const double* bc, const double* corr,
            const double* bCorr, double* flux, double* bFlux);
/* fvMatrix.C:594-740, fvMatrix<Type>::relax(alpha) with every patch
   non-coupled: bFaceCells and ic (nB x nCmpt) in patch order; diag
   (m's off-diagonals are used, diag is the in/out diagonal) and source
   (nCells x nCmpt) are updated in place */
void or_relax(const or_matrix* m, int32_t nB, const int32_t* bFaceCells, const double* ic,
              int32_t nCmpt, double alpha, const double* psi, double* diag, double* source);

/* ---- selection ids (names as in the reference run-time tables) ---- */
enum { OR_SOLVER_PCG = 0, OR_SOLVER_PBICGSTAB = 1, OR_SOLVER_PBICG = 2,
       OR_SOLVER_SMOOTH = 3, OR_SOLVER_GAMG = 4, OR_SOLVER_DIAGONAL = 5 };
enum { OR_PRECON_NONE = 0, OR_PRECON_DIAGONAL = 1, OR_PRECON_DIC = 2,
       OR_PRECON_FDIC = 3, OR_PRECON_DILU = 4, OR_PRECON_GAMG = 5 };
enum { OR_SMOOTH_GS = 0, OR_SMOOTH_SYMGS = 1, OR_SMOOTH_DIC = 2,
       OR_SMOOTH_DILU = 3, OR_SMOOTH_DICGS = 4, OR_SMOOTH_DILUGS = 5,
       OR_SMOOTH_FDIC = 6, OR_SMOOTH_NONE = 7 };

/* ---- preconditioners ---- */
/* rD setup: DICPreconditioner.C:57-84, DILUPreconditioner.C:57-85,
   FDICPreconditioner.C:46-83, diagonalPreconditioner.C:50-63 */
void or_calcReciprocalD_DIC(const or_matrix* m, double* rD);
void or_calcReciprocalD_DILU(const or_matrix* m, double* rD);
/* apply: w = M^-1 r */
void or_precondition(const or_matrix* m, int precon, const double* rD,
                     const double* r, double* w);
void or_preconditionT(const or_matrix* m, int precon, const double* rD,
                      const double* r, double* w);

/* ---- smoothers (lduMatrix/smoothers) ---- */
void or_smooth(const or_matrix* m, int smoother, const double* rD,
               double* psi, const double* source, int32_t nSweeps);

/* ---- solver controls / performance (lduMatrixSolver.C:158-164,
        SolverPerformance.H:78-268) ---- */
typedef struct or_controls_s
{
    int32_t solver;
    int32_t preconditioner;
    int32_t smoother;
    int32_t maxIter;       /* default 1000 */
    int32_t minIter;       /* default 0 */
    double  tolerance;     /* default 1e-6 */
    double  relTol;        /* default 0 */
    int32_t nSweeps;       /* smoothSolver, default 1 */
    /* GAMG (GAMGSolver.C:70-81, GAMGAgglomeration.C:248-257) */
    int32_t nPreSweeps, preSweepsLevelMultiplier, maxPreSweeps;
    int32_t nPostSweeps, postSweepsLevelMultiplier, maxPostSweeps;
    int32_t nFinestSweeps;
    int32_t scaleCorrection;      /* -1 = default (matrix.symmetric()) */
    int32_t directSolveCoarsest;
    int32_t nCellsInCoarsestLevel;
    int32_t mergeLevels;
    int32_t nVcycles;             /* GAMG preconditioner, default 2 */
    /* faceAreaPair agglomeration: the face weights of or_face_area_weights
       (NULL: algebraicPair, weights from the coefficients) */
    const double* faceWeights;
    /* GAMG preconditioner: the controls of its preconditioner
       sub-dictionary, which lduMatrix::preconditioner::New hands to
       GAMGPreconditioner (lduMatrixPreconditioner.C:72-82,
       GAMGPreconditioner.C:44-63); NULL: these controls */
    const struct or_controls_s* precondControls;
} or_controls;

/* faceAreaPairGAMGAgglomeration.C:55-80: w = mag(cmptMultiply(Sf/sqrt(magSf),
   (1, 1.01, 1.02))) per internal face; Sf [nFaces x 3] */
void or_face_area_weights(int32_t nFaces, const double* Sf, const double* magSf, double* w);

void or_controls_default(or_controls* c);

typedef struct
{
    double  initialResidual;
    double  finalResidual;
    int32_t nIterations;
    int32_t converged;
    int32_t singular;
    double  normFactor;
} or_perf;

/* Solve A psi = source in place.  history (optional, may be NULL) receives
   finalResidual after every iteration, history[0] = initial residual;
   it must hold maxIter+2 entries. */
int or_solve(const or_matrix* m, const or_controls* c, double* psi,
             const double* source, or_perf* perf, double* history);

/* pairGAMGAgglomeration::forward_ is a process-global toggle
   (pairGAMGAgglomeration.C:33); this resets it to true. */
void or_gamg_reset_direction(void);
int  or_gamg_get_direction(void);

/* ---- GAMG building blocks exposed for the tests ---- */
/* pairGAMGAgglomerate.C:123-301; returns nCoarse, fills restrictMap[nCells] */
int32_t or_pair_agglomerate(const or_addr* a, const double* faceWeights,
                            int32_t* restrictMap);

/* Full hierarchy (pairGAMGAgglomerate.C:31-118 + GAMGAgglomerateLduAddressing.C)
   nLevels = number of coarse levels.  Arrays owned by the struct. */
typedef struct
{
    int32_t  nLevels;
    int32_t* nCells;          /* [nLevels+1]: nCells[0] = finest */
    int32_t* nFaces;          /* [nLevels+1] */
    int32_t** lower;          /* [nLevels+1] addressing per level (level 0 borrowed) */
    int32_t** upper;
    int32_t** restrictMap;    /* [nLevels]: fine cell -> coarse cell */
    int32_t** faceRestrict;   /* [nLevels]: fine face -> coarse face or -1-cell */
    uint8_t** faceFlip;       /* [nLevels] */
} or_hierarchy;

int  or_gamg_agglomerate(const or_matrix* m, const or_controls* c,
                         or_hierarchy* h);
void or_hierarchy_free(or_hierarchy* h);

/* GAMGSolverAgglomerateMatrix.C:33-193, one level */
void or_agglomerate_matrix(int32_t nFineFaces, int32_t nCoarseCells,
                           int32_t nCoarseFaces, const int32_t* restrictMap,
                           int32_t nFineCells,
                           const int32_t* faceRestrict, const uint8_t* faceFlip,
                           const double* fineDiag, const double* fineUpper,
                           const double* fineLower /* NULL if symmetric */,
                           double* coarseDiag, double* coarseUpper,
                           double* coarseLower /* NULL if symmetric */);

/* scalarMatrices.C:77-180 and scalarMatricesTemplates.C:117-162 */
int  or_LUDecompose(int32_t n, double* A, int32_t* pivot);
void or_LUBacksubstitute(int32_t n, const double* LU, const int32_t* pivot,
                         double* x);

/* ---- finite-volume assembly (face loops feeding the solver) ---- */
/* gaussLaplacianScheme.C:45-89 + gaussLaplacianSchemes.C:43-56,
   surfaceInterpolationScheme.C:305-314 (linear interpolate),
   fixedValueFvPatchField.C:108-119, zeroGradientFvPatchField.C:138-155.
   Boundary faces are given flattened: bFaceCells, bMagSf, bDeltaCoeffs,
   bGamma, bType (0 zeroGradient, 1 fixedValue), bValue.  Outputs upper[F],
   diag[N] (negSumDiag), internalCoeffs[nB], boundaryCoeffs[nB]. */
void or_laplacian_assemble(const or_addr* a, const double* gammaCell,
                           const double* weights, const double* magSf,
                           const double* nonOrthDeltaCoeffs,
                           int32_t nB, const int32_t* bFaceCells,
                           const double* bMagSf, const double* bDeltaCoeffs,
                           const int32_t* bType, const double* bValue,
                           double* upper, double* diag,
                           double* internalCoeffs, double* boundaryCoeffs);

/* gaussConvectionScheme.C:119-170 with upwind weights pos0(phi)
   (upwind.H:116-128); outputs lower, upper, diag (negSumDiag) and patch
   coefficients. */
void or_upwind_assemble(const or_addr* a, const double* phi,
                        int32_t nB, const int32_t* bType, const double* bPhi,
                        const double* bValue,
                        double* lower, double* upper, double* diag,
                        double* internalCoeffs, double* boundaryCoeffs);

/* fvMatrix.C:112-175: diag += internalCoeffs, source += boundaryCoeffs */
void or_add_boundary(int32_t nB, const int32_t* bFaceCells,
                     const double* internalCoeffs, const double* boundaryCoeffs,
                     double* diag, double* source);

/* gaussGrad.C:30-75 with linear interpolation and
   gradScheme.C:123-148 boundary correction (the boundary gradient values
   are written to bGrad[nB*3]).  bPhi = boundary values of the field,
   bSnGrad = patch snGrad (0 for zeroGradient). Sf, bSf are [.*3]. */
void or_grad_gauss_linear(const or_addr* a, const double* phi,
                          const double* weights, const double* Sf,
                          const double* V,
                          int32_t nB, const int32_t* bFaceCells,
                          const double* bSf, const double* bMagSf,
                          const double* bPhi, const double* bSnGrad,
                          double* grad, double* bGrad);

/* ---- mesh geometry (primitiveMesh, surfaceInterpolation, fvPatch) ----
   Faces are polygons: face f has points facePoints[faceStart[f] ..
   faceStart[f+1]-1]; internal faces 0..nInternal-1 (owner lower[f],
   neighbour upper[f]) then boundary faces, owned by bFaceCells[b].
   Vectors are packed x, y, z. */
/* face::areaAndCentre (faceTemplates.C:79-149) per face, magSf =
   max(mag(Sf), rootVSmall) (primitiveMeshFaceCentresAndAreas.C) */
void or_face_geometry(int32_t nFaces, const int32_t* faceStart, const int32_t* facePoints,
                      const double* points, double* Cf, double* Sf, double* magSf);
/* primitiveMesh::makeCellCentresAndVols (primitiveMeshCellCentresAndVols.C:72-154);
   Cf/Sf of the internal faces then of the boundary faces */
void or_cell_geometry(int32_t nCells, int32_t nInternal, const int32_t* lower,
                      const int32_t* upper, int32_t nB, const int32_t* bFaceCells,
                      const double* Cf, const double* Sf, const double* bCf, const double* bSf,
                      double* C, double* V);
/* surfaceInterpolation::makeWeights / makeDeltaCoeffs /
   makeNonOrthDeltaCoeffs (surfaceInterpolation.C:135-330) for the internal
   faces, and the non-coupled patch values (fvPatch.C:159-164: delta =
   nf (nf & (Cf - Cn)); weights 1, deltaCoeffs 1/mag(delta),
   nonOrthDeltaCoeffs 1/max(nf & delta, 0.05 mag(delta))) */
void or_surface_interpolation(int32_t nInternal, const int32_t* lower, const int32_t* upper,
                              const double* C, const double* Cf, const double* Sf,
                              const double* magSf, double* weights, double* deltaCoeffs,
                              double* nonOrthDeltaCoeffs, int32_t nB,
                              const int32_t* bFaceCells, const double* bCf, const double* bSf,
                              const double* bMagSf, double* bDeltaCoeffs,
                              double* bNonOrthDeltaCoeffs);

#ifdef __cplusplus
}
#endif

#endif
