/*
 * foamhip.h -- C-ABI of the MI355X-native lduMatrix linear-solver library.
 *
 * The entry points replace, for one GPU, the run-time-selectable solver
 * stack that fvMatrix<scalar>::solveSegregated reaches through
 * lduMatrix::solver::New (src/OpenFOAM/matrices/lduMatrix/lduMatrix/
 * lduMatrixSolver.C:40-130).  A thin OpenFOAM adapter (hipPCG, hipDIC, ...)
 * maps its dictionary to fh_controls and its lduMatrix/lduAddressing arrays
 * to the plain pointers below.
 *
 * Conventions
 *   - every function returns FH_OK (0) or a non-zero FH_ERR_* status and never
 *     throws across the ABI; fh_last_error() returns the message of the last
 *     failure on the calling thread;
 *   - "host" arrays are caller-owned host memory; "device" arrays are
 *     caller-owned device memory on the handle's device;
 *   - label = int32_t, scalar = double (WM_LABEL_SIZE=32, WM_DP);
 *   - a handle is used by one host thread at a time; each fh_matrix owns one
 *     HIP stream, so independent matrices run concurrently.
 *   - coupled interfaces (processor/cyclic patches) are not part of this
 *     ABI: the matrices handed over are single-domain.
 */
#ifndef FOAMHIP_H
#define FOAMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FH_OK               0
#define FH_ERR_INVALID      1   /* bad argument / unknown name            */
#define FH_ERR_HIP          2   /* HIP runtime failure                     */
#define FH_ERR_NOMEM        3   /* device allocation failed               */
#define FH_ERR_UNSUPPORTED  4   /* valid request the library does not cover */
#define FH_ERR_NUMERIC      5   /* NaN/Inf in a residual or a setup factor */
#define FH_ERR_SCHEDULE     6   /* a device-side bounded wait expired      */

/* Selection ids.  Names accepted by the *_lookup functions are the
   reference run-time-selection names (PCG.C:34-35, PBiCGStab.C:34-38,
   DICPreconditioner.C:34-36, ...) with or without the "hip" prefix. */
enum
{
    FH_SOLVER_PCG = 0,
    FH_SOLVER_PBICGSTAB = 1,
    FH_SOLVER_PBICG = 2,
    FH_SOLVER_SMOOTH = 3,
    FH_SOLVER_GAMG = 4,
    FH_SOLVER_DIAGONAL = 5
};

enum
{
    FH_PRECON_NONE = 0,
    FH_PRECON_DIAGONAL = 1,
    FH_PRECON_DIC = 2,
    FH_PRECON_FDIC = 3,
    FH_PRECON_DILU = 4,
    FH_PRECON_GAMG = 5,
    /* no reference counterpart: DIC / DILU factorised on each 3-D tile of a
       hex box separately (block-Jacobi over tiles), used by the FP32 inner
       iterations of the mixed-precision path; on other meshes they fall back
       to DIC / DILU */
    FH_PRECON_TILE_DIC = 6,
    FH_PRECON_TILE_DILU = 7
};

enum
{
    FH_SMOOTH_GS = 0,
    FH_SMOOTH_SYMGS = 1,
    FH_SMOOTH_DIC = 2,
    FH_SMOOTH_DILU = 3,
    FH_SMOOTH_DICGS = 4,
    FH_SMOOTH_DILUGS = 5,
    FH_SMOOTH_FDIC = 6,
    FH_SMOOTH_NONE = 7
};

enum
{
    FH_PRECISION_DOUBLE = 0,  /* plain FP64 drop-in                        */
    FH_PRECISION_MIXED = 1    /* FP32 inner iterations, FP64 outer residual,
                                 dots, normFactor and refinement updates    */
};

/* Solver controls: the fvSolution entries read by
   lduMatrix::solver::readControls (lduMatrixSolver.C:158-164),
   smoothSolver::readControls (smoothSolver.C:70-74),
   GAMGSolver::readControls (GAMGSolver.C:351-388) and the agglomeration
   (GAMGAgglomeration.C:248-257, pairGAMGAgglomeration.C:44-52). */
typedef struct fh_controls
{
    int32_t solver;
    int32_t preconditioner;
    int32_t smoother;
    int32_t precision;
    int32_t maxIter;                 /* 1000 */
    int32_t minIter;                 /* 0    */
    double  tolerance;               /* 1e-6 */
    double  relTol;                  /* 0    */
    int32_t nSweeps;                 /* smoothSolver: 1 */
    int32_t nPreSweeps;              /* GAMG: 0 */
    int32_t preSweepsLevelMultiplier;/* 1 */
    int32_t maxPreSweeps;            /* 4 */
    int32_t nPostSweeps;             /* 2 */
    int32_t postSweepsLevelMultiplier;/* 1 */
    int32_t maxPostSweeps;           /* 4 */
    int32_t nFinestSweeps;           /* 2 */
    int32_t scaleCorrection;         /* -1: matrix.symmetric() */
    int32_t directSolveCoarsest;     /* 0 */
    int32_t nCellsInCoarsestLevel;   /* 10 */
    int32_t mergeLevels;             /* 1 */
    int32_t nVcycles;                /* GAMG preconditioner: 2 */
    /* mixed precision (no reference counterpart) */
    double  innerRelTol;             /* inner FP32 solve: residual reduction
                                        per refinement step, default 1e-4  */
    int32_t maxRefinements;          /* default 50 */
    int32_t innerPreconditioner;     /* mixed PCG: preconditioner of the FP32
                                        iterations; -1 = tile-local variant of
                                        `preconditioner` on hex boxes, else the
                                        same preconditioner                 */
    int32_t agglomerator;            /* GAMG: FH_AGGLOM_*, default
                                        algebraicPair (faceAreaPair, the
                                        reference's default, needs the mesh's
                                        face areas: fh_mesh_set_face_areas)  */
    /* GAMG preconditioner: the controls of the preconditioner
       sub-dictionary (preconditioner { preconditioner GAMG; smoother ...; }),
       which lduMatrix::preconditioner::New hands to GAMGPreconditioner as
       its whole dictionary (lduMatrixPreconditioner.C:72-82,
       GAMGPreconditioner.C:44-63): smoother, sweeps, agglomeration,
       nVcycles and the tolerance / relTol of its coarsest-level solve.
       NULL: the GAMG entries of these controls. */
    const struct fh_controls* preconditionerControls;
} fh_controls;

/* GAMGAgglomeration types (GAMGAgglomeration.C:291-340) */
enum
{
    FH_AGGLOM_ALGEBRAIC_PAIR = 0,   /* algebraicPairGAMGAgglomeration.C:57-64:
                                       face weights from the coefficients  */
    FH_AGGLOM_FACE_AREA_PAIR = 1    /* faceAreaPairGAMGAgglomeration.C:55-80:
                                       face weights from the face areas   */
};

/* SolverPerformance<scalar> (SolverPerformance.H:84-92) plus extras. */
typedef struct fh_perf
{
    char    solverName[64];          /* e.g. "DICPCG" (PCG.C:73-77 naming) */
    double  initialResidual;
    double  finalResidual;
    int32_t nIterations;             /* inner iterations in mixed mode      */
    int32_t converged;
    int32_t singular;
    int32_t nRefinements;            /* mixed mode outer steps (0 in FP64)  */
    double  normFactor;
    double  solveMilliseconds;       /* device time of the solve           */
} fh_perf;

typedef struct fh_mesh fh_mesh;
typedef struct fh_matrix fh_matrix;

typedef struct fh_mesh_info
{
    int32_t nCells;
    int32_t nFaces;
    int32_t nLevels;                 /* depth of the lower-triangular DAG  */
    int32_t nParts;                  /* pipelined sweep partitions (0=none)*/
    int32_t boxNx, boxNy, boxNz;     /* detected hex-box shape, 0 if not   */
    int32_t tiled;                   /* 1: hex box swept in 3-D tiles      */
    int64_t deviceBytes;             /* addressing bytes resident in HBM   */
} fh_mesh_info;

/* ---- library ---------------------------------------------------------- */
int         fh_version(void);
const char* fh_last_error(void);
int         fh_set_device(int32_t device);
void        fh_controls_default(fh_controls* c);
int         fh_solver_lookup(const char* name, int32_t* id);
int         fh_preconditioner_lookup(const char* name, int32_t* id);
int         fh_smoother_lookup(const char* name, int32_t* id);
/* pairGAMGAgglomeration::forward_ (pairGAMGAgglomeration.C:33) is a
   process-global toggle in the reference; the library keeps the same
   state.  These read and reset it (reset = true). */
int         fh_gamg_direction(void);
void        fh_gamg_reset_direction(void);
/* The pair-agglomeration hierarchy a GAMG solve of this matrix with these
   controls uses (built on the device on first use and cached with the mesh,
   as GAMGAgglomeration is cached on the mesh database): the number of
   levels (GAMGAgglomeration::size, GAMGAgglomeration.H:325), and per
   level sizes[4] = {fine cells, fine faces, coarse cells, coarse faces}
   and, for each non-NULL host array, restrictAddressing [fine cells],
   faceRestrictAddressing [fine faces] (coarse face, or -1 - coarse cell
   for a face inside a coarse cell), faceFlipMap [fine faces] and the
   coarse level's lduAddressing lower/upper [coarse faces]
   (GAMGAgglomeration.H:331-367), in the caller's cell and face order. */
int fh_gamg_agglomeration_levels(fh_matrix* matrix, const fh_controls* controls,
                                 int32_t* nLevels);
int fh_gamg_agglomeration_level(fh_matrix* matrix, const fh_controls* controls,
                                int32_t level, int32_t* sizes, int32_t* restrictMap,
                                int32_t* faceRestrict, uint8_t* faceFlip, int32_t* lower,
                                int32_t* upper);

/* ---- mesh: lduAddressing (lduAddressing.H:112-207) --------------------
   lowerAddr = owner, upperAddr = neighbour of each internal face, faces in
   upper-triangular order.  Builds the device row structure and the sweep
   schedule once per mesh. */
int fh_mesh_create(int32_t nCells, int32_t nFaces, const int32_t* lowerAddr,
                   const int32_t* upperAddr, fh_mesh** mesh);
int fh_mesh_info_get(const fh_mesh* mesh, fh_mesh_info* info);
/* The face area vectors Sf [nFaces x 3] and magnitudes magSf [nFaces] of
   the internal faces (fvMesh::Sf(), magSf()), host arrays: the weights of
   the faceAreaPair agglomeration, |(Sf/sqrt(magSf)) * (1, 1.01, 1.02)|
   (faceAreaPairGAMGAgglomeration.C:55-80).  A change drops the mesh's
   cached agglomeration. */
int fh_mesh_set_face_areas(fh_mesh* mesh, const double* Sf, const double* magSf);
/* A mesh with coupled (cyclic) interfaces, the lduInterfaceFieldPtrsList
   of the matrix (lduMatrixUpdateMatrixInterfaces.C:30-268, one GPU: the
   patches' cells are all on this device).  Interface i owns faces
   interfaceStart[i] .. interfaceStart[i+1]-1 of faceCells / nbrFaceCells
   (concatenated in interface order): face e couples faceCells[e] to
   nbrFaceCells[e], as cyclicFvPatchField::updateInterfaceMatrix does with
   the neighbour patch's faceCells (cyclicFvPatchField.C:150-174; scalar
   fields, no transformation).  Amul, Tmul, residual, sumA, normFactor and
   the Gauss-Seidel smoothers add the interfaces after the face loop in this
   order; DIC/DILU ignore them, as the reference's do.  GAMG on a coupled
   mesh is FH_ERR_UNSUPPORTED.  nInterfaces 0 is fh_mesh_create. */
int fh_mesh_create_coupled(int32_t nCells, int32_t nFaces, const int32_t* lowerAddr,
                           const int32_t* upperAddr, int32_t nInterfaces,
                           const int32_t* interfaceStart, const int32_t* faceCells,
                           const int32_t* nbrFaceCells, fh_mesh** mesh);
int fh_mesh_destroy(fh_mesh* mesh);

/* ---- matrix: lduMatrix coefficients (lduMatrix.H:79-87) ----------------
   upper == NULL: diagonal matrix; lower == NULL: symmetric (lower aliases
   upper, lduMatrix.C:168-312).  Host arrays in the caller's face order. */
int fh_matrix_create(fh_mesh* mesh, fh_matrix** matrix);
int fh_matrix_set(fh_matrix* matrix, const double* diag, const double* upper,
                  const double* lower);
int fh_matrix_set_device(fh_matrix* matrix, const double* d_diag,
                         const double* d_upper, const double* d_lower);
/* interfaceBouCoeffs and interfaceIntCoeffs of a coupled mesh's matrix
   (host arrays, one value per interface face in the mesh's interface-face
   order; intCoeffs NULL: the bouCoeffs), as fvMatrix::solveSegregated hands
   them to the solver (fvMatrixSolve.C:103-216); required before a solve on
   a coupled mesh, and kept across fh_matrix_set calls. */
int fh_matrix_set_interfaces(fh_matrix* matrix, const double* interfaceBouCoeffs,
                             const double* interfaceIntCoeffs);
/* lduMatrix::diag(), upper(), lower() (lduMatrix.C:168-312): copies of the
   coefficients, natural order; any output may be NULL.  lower of a
   symmetric matrix is its upper (lower() returns upper when lowerPtr_ is
   unset); the off-diagonals of a diagonal-only matrix are zeros. */
int fh_matrix_get(fh_matrix* matrix, double* diag, double* upper, double* lower);
int fh_matrix_destroy(fh_matrix* matrix);

/* ---- solve: lduMatrix::solver::solve (lduMatrix.H:250-255) -------------
   psi holds the initial guess on entry and the solution on return. */
int fh_solve(fh_matrix* matrix, const fh_controls* controls, double* psi,
             const double* source, fh_perf* perf);
int fh_solve_device(fh_matrix* matrix, const fh_controls* controls,
                    double* d_psi, const double* d_source, fh_perf* perf);

/* ---- single operations (host arrays), the lduMatrix members -------------*/
/* lduMatrixATmul.C:34-92 */
int fh_amul(fh_matrix* m, const double* psi, double* Apsi);
/* lduMatrixATmul.C:95-151 */
int fh_tmul(fh_matrix* m, const double* psi, double* Tpsi);
/* lduMatrixATmul.C:203-280 */
int fh_residual(fh_matrix* m, const double* psi, const double* source,
                double* rA);
/* lduMatrixATmul.C:154-200 */
int fh_sum_a(fh_matrix* m, double* sumA);
/* The off-diagonal face loops around the solve (pressure-velocity
   coupling).  A field of a Type with nCmpt components (1..9: scalar,
   vector, symmTensor, tensor ...) holds nCells (or nFaces) x nCmpt values,
   component fastest: the memory layout of Field<Type>.  A diagonal-only
   matrix gives H = 0 and H1 = 0 and leaves sumOff unchanged. */
/* lduMatrixTemplates.C:33-64: Hpsi = -(off-diagonal part of A) psi */
int fh_h(fh_matrix* m, int32_t nCmpt, const double* psi, double* Hpsi);
/* lduMatrixATmul.C:298-324: H1 = -(sum of the row's off-diagonal
   coefficients) */
int fh_h1(fh_matrix* m, double* H1);
/* lduMatrixTemplates.C:77-111: faceHpsi[f] = upper[f] psi[u[f]]
   - lower[f] psi[l[f]], natural face order; FH_ERR_INVALID for a
   diagonal-only matrix (a fatal error in the reference) */
int fh_face_h(fh_matrix* m, int32_t nCmpt, const double* psi, double* faceHpsi);
/* lduMatrixOperations.C:77-97: sumOff += the row's sum of |off-diagonal
   coefficient| (sumOff is read and written) */
int fh_sum_mag_off_diag(fh_matrix* m, double* sumOff);
/* fh_h and fh_face_h on device arrays (natural order, component fastest),
   for callers whose fields stay in HBM */
int fh_h_device(fh_matrix* m, int32_t nCmpt, const double* d_psi, double* d_Hpsi);
int fh_face_h_device(fh_matrix* m, int32_t nCmpt, const double* d_psi, double* d_faceHpsi);
/* lduMatrixSolver.C:174-197 */
int fh_norm_factor(fh_matrix* m, const double* psi, const double* source,
                   double* normFactor);
/* No counterpart in the reference: fh_norm_factor for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  psi and source are device
   arrays in the caller's cell order, normFactor one device double; the
   result is written on the matrix's stream (synchronize before reading it)
   and equals fh_norm_factor's bits.  It calls no hipMalloc, hipFree or
   allocating routine: its reduction state lives in a static device array
   of the library (about 1 MiB per device, held until the process ends),
   so meshes are limited to 2^24 cells (FH_ERR_INVALID above) and calls that
   overlap on one device race. */
int fh_norm_factor_device_noalloc(fh_matrix* m, const double* d_psi, const double* d_source,
                                  double* d_normFactor);
/* DICPreconditioner.C:57-84 / DILUPreconditioner.C:57-85 /
   diagonalPreconditioner.C:50-63: the reciprocal factor rD */
int fh_reciprocal_d(fh_matrix* m, int32_t preconditioner, double* rD);
/* lduMatrix::preconditioner::precondition / preconditionT
   (lduMatrix.H:468-487) */
int fh_precondition(fh_matrix* m, int32_t preconditioner, const double* rA,
                    double* wA);
int fh_precondition_t(fh_matrix* m, int32_t preconditioner, const double* rA,
                      double* wA);
/* lduMatrix::smoother::smooth (lduMatrix.H:399-405) */
int fh_smooth(fh_matrix* m, int32_t smoother, double* psi,
              const double* source, int32_t nSweeps);
/* The FP32 preconditioner application of the mixed-precision inner
   iterations (host float arrays, natural order); no reference counterpart,
   it is the single-precision form of fh_precondition. */
int fh_precondition_f32(fh_matrix* m, int32_t preconditioner, const float* rA, float* wA);

/* ---- measurement (no reference counterpart) -----------------------------
   Times `reps` back-to-back launches of one kernel of the solve on the
   matrix's stream with HIP events.  kernel: FH_KERNEL_*; precision:
   FH_PRECISION_DOUBLE or FH_PRECISION_MIXED (FP32 values).  Returns the
   average milliseconds per launch and the algorithmic HBM bytes one launch
   moves (each array read or written once, indices int32). */
enum
{
    FH_KERNEL_SPMV = 0,          /* y = A x                         */
    FH_KERNEL_SWEEP_FORWARD = 1, /* DIC/DILU forward substitution    */
    FH_KERNEL_SWEEP_BACKWARD = 2,/* DIC/DILU backward substitution   */
    FH_KERNEL_AXPY_NORM = 3,     /* psi += a p, r -= a w, sum |r|     */
    FH_KERNEL_PRECONDITION = 4,  /* the whole preconditioner apply of the
                                    PCG loop, with its fused w.r sum    */
    FH_KERNEL_H = 5              /* H x of a scalar field (fh_h), FP64   */
};
int fh_profile_kernel(fh_matrix* matrix, int32_t kernel, int32_t precision,
                      int32_t preconditioner, int32_t reps, double* msPerLaunch,
                      double* bytesPerLaunch);
/* Live timing inside the PCG loops (no reference counterpart): with bit
   (1 << FH_KERNEL_*) set in kernelMask, every launch of that kernel in the
   following solves is bracketed by HIP events on the matrix's stream and
   the launches of the iterations that ran are summed (PRECONDITION, SPMV,
   AXPY_NORM of the box PCG loops).  A call resets the sums; mask 0 stops. */
int fh_kernel_timing(fh_matrix* matrix, uint32_t kernelMask);
int fh_kernel_timing_get(fh_matrix* matrix, int32_t kernel, double* msTotal,
                         int64_t* launches);
/* Streaming bandwidth of the current device (no reference counterpart; the
   achievable figure beside the published peak): a grid-stride kernel over
   bytes-sized arrays with 16-byte accesses, best of five launches, in
   bytes moved per second.  pattern: FH_STREAM_READ (one array read),
   FH_STREAM_COPY (read one, write one), FH_STREAM_3R1W (read three, write
   one: the mix of the box product). */
enum
{
    FH_STREAM_READ = 0,
    FH_STREAM_COPY = 1,
    FH_STREAM_3R1W = 2
};
int fh_profile_stream(int32_t pattern, int64_t bytes, double* bytesPerSecond);

/* ---- finite-volume assembly on the device ------------------------------
   Boundary faces of all non-coupled patches, flattened in patch order
   (fvBoundaryMesh order); bFaceCells = fvPatch::faceCells.  Device arrays
   are in the caller's (natural) cell and face order. */
typedef struct fh_boundary fh_boundary;
int fh_boundary_create(fh_mesh* mesh, int32_t nBFaces, const int32_t* bFaceCells,
                       fh_boundary** boundary);
int fh_boundary_destroy(fh_boundary* boundary);

/* fvm::laplacian(gamma, psi), Gauss linear uncorrected
   (gaussLaplacianScheme.C:45-89, gaussLaplacianSchemes.C:43-56): writes the
   matrix coefficients (symmetric), the patch internal/boundary coefficients
   (bType 1 fixedValue with bValue, 0 zeroGradient).  addBoundaryDiag != 0
   includes the internal coefficients in the diagonal and d_source != NULL
   receives the boundary coefficients, as fvMatrix::solveSegregated does
   before the solve (fvScalarMatrix.C:155-159). */
int fh_assemble_laplacian(fh_matrix* matrix, const fh_boundary* boundary,
                          const double* d_gamma, const double* d_weights,
                          const double* d_magSf, const double* d_deltaCoeffs,
                          const double* d_bMagSf, const double* d_bDeltaCoeffs,
                          const int32_t* d_bType, const double* d_bValue,
                          double* d_internalCoeffs, double* d_boundaryCoeffs,
                          double* d_source, int32_t addBoundaryDiag);

/* fvm::div(phi, psi) with upwind weights minus fvm::laplacian(D, psi)
   (gaussConvectionScheme.C:119-170, upwind.H:116-128) */
int fh_assemble_convection_diffusion(fh_matrix* matrix, const fh_boundary* boundary,
                                     const double* d_phi, double D,
                                     const double* d_magSf, const double* d_deltaCoeffs,
                                     const double* d_bPhi, const double* d_bMagSf,
                                     const double* d_bDeltaCoeffs, const int32_t* d_bType,
                                     const double* d_bValue, double* d_internalCoeffs,
                                     double* d_boundaryCoeffs, double* d_source,
                                     int32_t addBoundaryDiag);

/* fvMatrix<Type>::relax(alpha) (fvMatrix.C:594-740) for a matrix whose
   patches are all non-coupled (those of fh_boundary_create): with D0 the
   diagonal, D = D0 plus each boundary face's largest internal-coefficient
   magnitude (patch order), D = max(|D|, sumMagOffDiag)/alpha, less each
   face's smallest internal coefficient; source += (D - D0) psi.  The
   matrix's diagonal becomes D (cached factors are dropped).  Device arrays,
   natural order, component fastest: d_internalCoeffs nBFaces x nCmpt,
   d_psi and d_source nCells x nCmpt (d_source updated in place); d_diag
   (may be NULL) receives D.  alpha <= 0 does nothing, as in the
   reference. */
int fh_relax(fh_matrix* matrix, const fh_boundary* boundary, double alpha, int32_t nCmpt,
             const double* d_internalCoeffs, const double* d_psi, double* d_source,
             double* d_diag);

/* fvMatrix<Type>::flux() (fvMatrix.C:953-1028) for a matrix whose patches
   are all non-coupled: internal faces upper psi(neighbour) - lower
   psi(owner) (faceH, per component), boundary faces internalCoeffs *
   psi(faceCell) - boundaryCoeffs (component by component), each plus the
   face flux correction when d_corr / d_bCorr are given (NULL: none).
   Device arrays, natural order, component fastest: d_psi nCells x nCmpt,
   d_internalCoeffs / d_boundaryCoeffs / d_bCorr / d_bFlux nBFaces x nCmpt,
   d_corr / d_flux nFaces x nCmpt. */
int fh_flux(fh_matrix* matrix, const fh_boundary* boundary, int32_t nCmpt,
            const double* d_psi, const double* d_internalCoeffs,
            const double* d_boundaryCoeffs, const double* d_corr, const double* d_bCorr,
            double* d_flux, double* d_bFlux);

/* fvc::grad(phi), Gauss linear (gaussGrad.C:30-75) with the patch-normal
   boundary correction (gradScheme.C:123-148).  d_grad [nCells*3],
   d_bGrad [nBFaces*3] (may be NULL). */
int fh_grad_gauss_linear(fh_mesh* mesh, const fh_boundary* boundary, const double* d_phi,
                         const double* d_weights, const double* d_Sf, const double* d_V,
                         const double* d_bSf, const double* d_bMagSf, const double* d_bPhi,
                         const double* d_bSnGrad, double* d_grad, double* d_bGrad);

/* ---- mesh geometry (device arrays; vectors packed x, y, z) -------------
   Faces are polygons: face f has the points facePoints[faceStart[f] ..
   faceStart[f+1]-1], internal faces first (in the mesh's face order), then
   the boundary faces (in the boundary's bFaceCells order). */
/* face::areaAndCentre per face (faceTemplates.C:79-149) and magSf =
   max(mag(Sf), rootVSmall) (primitiveMeshFaceCentresAndAreas.C):
   Cf, Sf [3*nFaces], magSf [nFaces] */
int fh_face_geometry(int32_t nFaces, const int32_t* d_faceStart, const int32_t* d_facePoints,
                     const double* d_points, double* d_Cf, double* d_Sf, double* d_magSf);
/* primitiveMesh::makeCellCentresAndVols (primitiveMeshCellCentresAndVols.C:72-154):
   C [3*nCells], V [nCells] from the internal (Cf, Sf) and boundary (bCf,
   bSf) face centres and areas */
int fh_cell_geometry(fh_mesh* mesh, const fh_boundary* boundary, const double* d_Cf,
                     const double* d_Sf, const double* d_bCf, const double* d_bSf,
                     double* d_C, double* d_V);
/* surfaceInterpolation::makeWeights, makeDeltaCoeffs, makeNonOrthDeltaCoeffs
   (surfaceInterpolation.C:135-330) of the internal faces, and of the
   non-coupled patch faces deltaCoeffs = 1/mag(delta) and
   nonOrthDeltaCoeffs = 1/max(nf & delta, 0.05 mag(delta)) with delta =
   nf (nf & (Cf - Cn)) (fvPatch.C:138-164); d_bNonOrthDeltaCoeffs may be
   NULL */
int fh_surface_interpolation(fh_mesh* mesh, const fh_boundary* boundary, const double* d_C,
                             const double* d_Cf, const double* d_Sf, const double* d_magSf,
                             const double* d_bCf, const double* d_bSf, const double* d_bMagSf,
                             double* d_weights, double* d_deltaCoeffs,
                             double* d_nonOrthDeltaCoeffs, double* d_bDeltaCoeffs,
                             double* d_bNonOrthDeltaCoeffs);

/* ------------------------------------------------------------------------
   Serving: a request loop for many independent solves.  No counterpart in
   the reference (which solves one system per call); it runs the fh_solve
   path above on several HIP streams at once.  Requests are copied at
   submission, admitted in arrival order as a solve slot and device memory
   (an estimate of the request's footprint against memoryBudget) become
   free, and solved on resident matrices kept per mesh, so meshes, factors
   and GAMG hierarchies stay in HBM between requests.  A result is the same
   bits whether the request ran alone or beside others.
   ------------------------------------------------------------------------ */
typedef struct fh_server fh_server;

typedef struct fh_server_config
{
    int32_t maxConcurrent;      /* solves in flight, one stream each (default 8) */
    int32_t borrowInputs;       /* 0: copy the request arrays at submission (default);
                                   1: use them in place -- the caller keeps them valid
                                   and unchanged until fh_server_wait returns the result */
    int64_t memoryBudget;       /* device bytes of the resident matrices (default 32 GiB) */
} fh_server_config;

typedef struct fh_request
{
    int64_t id;                 /* caller's request id (unique among pending ones) */
    fh_mesh* mesh;              /* caller-owned, outlives the request */
    const double* diag;         /* host arrays (copied unless borrowInputs) */
    const double* upper;
    const double* lower;        /* NULL: symmetric */
    const double* source;
    const double* psi0;         /* NULL: zero initial guess */
    fh_controls controls;
} fh_request;

typedef struct fh_result
{
    int64_t id;
    int32_t status;             /* FH_OK or the solve's error status */
    int32_t nCells;
    fh_perf perf;
    double submitTime;          /* server clock, seconds */
    double startTime;
    double endTime;
    char message[160];          /* error text when status != FH_OK */
} fh_result;

void fh_server_config_default(fh_server_config* config);
int  fh_server_create(const fh_server_config* config, fh_server** out);
/* queues the request (copying its arrays unless borrowInputs); returns at once */
int  fh_server_submit(fh_server* server, const fh_request* request);
/* next completed result (completion order); *got = 0 after timeoutSeconds */
int  fh_server_wait(fh_server* server, double timeoutSeconds, fh_result* result,
                    int32_t* got);
/* copy the solution of a completed request out and release it */
int  fh_server_fetch(fh_server* server, int64_t id, double* psi);
/* seconds since the server was created (the clock of fh_result) */
int  fh_server_now(fh_server* server, double* seconds);
/* requests submitted and not yet returned by fh_server_wait */
int  fh_server_pending(fh_server* server, int64_t* count);
int  fh_server_destroy(fh_server* server);

#ifdef __cplusplus
}
#endif

#endif
