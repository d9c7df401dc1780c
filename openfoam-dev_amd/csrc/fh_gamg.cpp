// GAMG: agglomeration hierarchy, Galerkin coarse matrices and the V-cycle
// (GAMGSolver.C:48-332, GAMGSolverSolve.C:31-552, GAMGSolverScale.C:31-76,
// GAMGSolverAgglomerateMatrix.C:33-193, GAMGPreconditioner.C:87-148).
//
// The pairwise agglomeration (pairGAMGAgglomerate.C:31-301, fh_agglom.hip)
// runs on the device once per mesh (the reference caches it on the mesh
// database, GAMGAgglomeration.C:357-399), as does everything that runs per
// solve -- coarse coefficients, restriction, prolongation, scaling,
// smoothing and the coarsest solve.  The coarse levels' row structures and
// sweep schedules are built on the host from the downloaded addressing.
#include "fh_agglom.h"
#include "fh_gamg.h"
#include "fh_gamg_kernels.h"


#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

namespace fh
{

// pairGAMGAgglomeration::forward_ (pairGAMGAgglomeration.C:33)
bool g_forward = true;

// The coarse addressing of one hierarchy level, downloaded for the coarse
// mesh build
struct HostLevel
{
    int32_t nFineCells, nCoarseCells;
    std::vector<int32_t> cLower, cUpper;
};

template<class T>
static void download(const DBuf<T>& d, size_t n, std::vector<T>& h, hipStream_t s)
{
    h.resize(n);
    if (n) FH_HIP(hipMemcpyAsync(h.data(), d.p, n*sizeof(T), hipMemcpyDeviceToHost, s));
}

// ---------------------------------------------------------------------------
// device agglomeration (per mesh)
// ---------------------------------------------------------------------------

static fh_mesh* make_mesh(int device, int32_t n, const std::vector<int32_t>& lo,
                          const std::vector<int32_t>& up)
{
    fh_mesh* m = new fh_mesh();
    m->device = device;
    try
    {
        FH_HIP(hipStreamCreateWithFlags(&m->setupStream, hipStreamNonBlocking));
        build_mesh(m, n, (int32_t)lo.size(), lo.data(), up.data());
        // FOAMHIP_GAMG_VERBOSE: the shape of each coarse level and how its
        // sweeps run
        static const bool verbose = std::getenv("FOAMHIP_GAMG_VERBOSE") != nullptr;
        if (verbose)
            std::fprintf(stderr, "foamhip GAMG level: %d cells, %d faces, %d dependency levels, "
                                 "%s\n", n, (int)lo.size(), m->sched.nLevels,
                         m->chunkSweep ? "chunk sweep" : "one-workgroup level sweep");
    }
    catch (...)
    {
        if (m->setupStream) (void)hipStreamDestroy(m->setupStream);
        delete m;
        throw;
    }
    return m;
}

// a level: its maps, its coarse mesh and the transfer lists (rStart/rCells,
// prolong, iStart/iFaces, fStart/fFaces/fFlip) in library row numbering
struct AggLevel : TransferLists
{
    std::unique_ptr<AggMaps> maps;      // the level's maps in natural numbering
    fh_mesh* mesh = nullptr;            // the coarse mesh of this level
    int32_t nFine = 0, nCoarse = 0, nCoarseFaces = 0;

    ~AggLevel()
    {
        if (mesh)
        {
            if (mesh->setupStream) (void)hipStreamDestroy(mesh->setupStream);
            delete mesh;
        }
    }
};

struct GamgAgglomeration
{
    int32_t minCells = -1, mergeLevels = -1, kind = -1;
    std::vector<std::unique_ptr<AggLevel>> levels;
    // natural rows of the coarsest mesh for the single-workgroup solve
    DBuf<int32_t> csRowStart, csEnt, csCol;
};

static std::shared_ptr<GamgAgglomeration> build_agglomeration(fh_matrix* M,
                                                              const fh_controls* c)
{
    fh_mesh* mesh = M->mesh;
    const int32_t F = mesh->nFaces;
    hipStream_t s = M->stream;
    // algebraicPairGAMGAgglomeration.C:57-64: face weights mag(upper), or
    // max(mag(upper), mag(lower)) for an asymmetric matrix;
    // faceAreaPairGAMGAgglomeration.C:55-80: the mesh's face-area weights
    const bool byArea = c->agglomerator == FH_AGGLOM_FACE_AREA_PAIR;
    if (byArea && (int32_t)mesh->faceAreaWeights.size() != F)
        fail(FH_ERR_INVALID, "agglomerator faceAreaPair needs the mesh's face areas "
                             "(fh_mesh_set_face_areas)");
    DBuf<double> w;
    w.alloc(std::max(F, 1));
    if (byArea)
        w.upload(mesh->faceAreaWeights, s);
    else
        launch_agg_weights(F, M->d_upperF.p, M->symmetric ? nullptr : M->d_lowerF.p, w.p, s);
    // FOAMHIP_GAMG_VERBOSE: the time of each step of the hierarchy build
    static const bool verbose = std::getenv("FOAMHIP_GAMG_VERBOSE") != nullptr;
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    auto t0 = clk::now();
    std::vector<HostLevel> host;
    auto dev = device_agglomerate(mesh->nCells, F, mesh->d_lower.p, mesh->d_upper.p, w.p,
                                  c->nCellsInCoarsestLevel, c->mergeLevels, g_forward, s);
    {
        for (const auto& D : dev)
        {
            HostLevel L;
            L.nFineCells = D->nFine;
            L.nCoarseCells = D->nCoarse;
            download(D->lower, D->nCoarseFaces, L.cLower, s);
            download(D->upper, D->nCoarseFaces, L.cUpper, s);
            host.push_back(std::move(L));
        }
        FH_HIP(hipStreamSynchronize(s));
    }
    if (verbose)
        std::fprintf(stderr, "foamhip GAMG agglomeration: %zu levels in %.1f ms\n",
                     host.size(), ms(t0));
    auto A = std::make_shared<GamgAgglomeration>();
    A->minCells = c->nCellsInCoarsestLevel;
    A->mergeLevels = c->mergeLevels;
    A->kind = c->agglomerator;
    const fh_mesh* fm = mesh;
    for (size_t li = 0; li < host.size(); li++)
    {
        const HostLevel& L = host[li];
        auto D = std::make_unique<AggLevel>();
        D->maps = std::move(dev[li]);
        t0 = clk::now();
        D->mesh = make_mesh(mesh->device, L.nCoarseCells, L.cLower, L.cUpper);
        const double meshMs = ms(t0);
        t0 = clk::now();
        const fh_mesh* cm = D->mesh;
        D->nFine = L.nFineCells;
        D->nCoarse = L.nCoarseCells;
        D->nCoarseFaces = (int32_t)L.cLower.size();
        hipStream_t us = cm->setupStream;
        // restriction / Galerkin diagonal lists, prolongation, intra-cluster
        // and coarse-face fine faces (fh_agglom.hip)
        device_transfer_lists(*D->maps, fm->d_perm.p, cm->d_perm.p, *D, us);
        FH_HIP(hipStreamSynchronize(us));
        if (verbose)
            std::fprintf(stderr, "foamhip GAMG level %zu: mesh %.1f ms, transfer lists %.1f ms\n",
                         A->levels.size(), meshMs, ms(t0));
        fm = cm;
        A->levels.push_back(std::move(D));
    }
    if (!A->levels.empty())
    {
        // natural rows of the coarsest mesh: lower entries then upper entries,
        // each in face order (the per-row order of lduMatrix::Amul)
        const fh_mesh* cm = A->levels.back()->mesh;
        const int32_t n = cm->nCells, nF = cm->nFaces;
        std::vector<int32_t> rs(n + 1, 0), ent(2*(size_t)nF), col(2*(size_t)nF);
        for (int32_t f = 0; f < nF; f++) { rs[cm->upper[f] + 1]++; rs[cm->lower[f] + 1]++; }
        for (int32_t i = 0; i < n; i++) rs[i + 1] += rs[i];
        std::vector<int32_t> fill(rs.begin(), rs.end() - 1);
        for (int32_t f = 0; f < nF; f++)
        {
            const int32_t u = cm->upper[f];
            ent[fill[u]] = f << 1;
            col[fill[u]++] = cm->lower[f];
        }
        for (int32_t f = 0; f < nF; f++)
        {
            const int32_t l = cm->lower[f];
            ent[fill[l]] = (f << 1) | 1;
            col[fill[l]++] = cm->upper[f];
        }
        A->csRowStart.upload(rs, s);
        A->csEnt.upload(ent, s);
        A->csCol.upload(col, s);
        FH_HIP(hipStreamSynchronize(s));
    }
    return A;
}

// ---------------------------------------------------------------------------
// per-matrix hierarchy: coarse matrices, coarsest solve data, work vectors
// ---------------------------------------------------------------------------

template<class T>
struct GamgWork
{
    std::vector<std::unique_ptr<DBuf<T>>> corr, src;   // per coarse level
    DBuf<T> s1, s2;                                    // level scratch (largest level)
    DBuf<T> Apsi, fCorr, fRes;                         // finest level
    DBuf<T> e, r;                                      // mixed solver correction / residual
    bool ready = false;
};

enum { COARSEST_ITERATIVE = 0, COARSEST_LU = 1, COARSEST_INVERSE = 2 };

// LU back-substitution is one lane: it is used up to this size, the dense
// explicit inverse (GEMMs) above it
static const int32_t kLuMaxCells = 96;
static const int32_t kInverseMaxCells = 4096;

struct GamgHierarchy
{
    std::shared_ptr<GamgAgglomeration> agg;
    std::vector<fh_matrix*> mats;        // coarse level matrices, on the parent's stream
    bool coeffsValid = false;
    fh_controls ctl{};
    bool scale = false;
    int coarseMode = -1;
    bool coarseValid = false;
    DBuf<double> csDiagN, csRD, csSumA, csWork;
    DBuf<double> luLU, luWork;
    DBuf<int32_t> luPivot;
    DBuf<double> dA, dX, dX1, dT, dPart, dGemv;
    DBuf<NsState> dNs;
    DBuf<int32_t> dFlag;
    GamgWork<double> w64;
    GamgWork<float> w32;

    ~GamgHierarchy()
    {
        for (fh_matrix* A : mats)
        {
            matrix_free(A);
            delete A;
        }
    }
};

template<class T> static GamgWork<T>& work_of(GamgHierarchy* H);
template<> GamgWork<double>& work_of<double>(GamgHierarchy* H) { return H->w64; }
template<> GamgWork<float>& work_of<float>(GamgHierarchy* H) { return H->w32; }

template<class T>
static GamgWork<T>& ensure_work(fh_matrix* M, GamgHierarchy* H)
{
    GamgWork<T>& W = work_of<T>(H);
    if (W.ready) return W;
    const int32_t N = M->mesh->nCells;
    int32_t maxN = 1;
    W.corr.clear();
    W.src.clear();
    for (fh_matrix* A : H->mats)
    {
        const int32_t n = A->mesh->nCells;
        maxN = std::max(maxN, n);
        W.corr.push_back(std::make_unique<DBuf<T>>());
        W.src.push_back(std::make_unique<DBuf<T>>());
        W.corr.back()->alloc(n);
        W.src.back()->alloc(n);
    }
    W.s1.alloc(maxN);
    W.s2.alloc(maxN);
    W.Apsi.alloc(N);
    W.fCorr.alloc(N);
    W.fRes.alloc(N);
    W.e.alloc(N);
    W.r.alloc(N);
    W.ready = true;
    return W;
}

void gamg_invalidate(fh_matrix* M)
{
    GamgHierarchy* H = static_cast<GamgHierarchy*>(M->gamg);
    if (!H) return;
    H->coeffsValid = false;
    H->coarseValid = false;
}

void gamg_release(fh_matrix* M)
{
    delete static_cast<GamgHierarchy*>(M->gamg);
    M->gamg = nullptr;
}

void gamg_check_sweep_errors(fh_matrix* M)
{
    GamgHierarchy* H = static_cast<GamgHierarchy*>(M->gamg);
    if (!H) return;
    for (fh_matrix* A : H->mats) check_sweep_errors(A);
}

bool gamg_forward() { return g_forward; }
void gamg_set_forward(bool f) { g_forward = f; }

static void setup_coarsest(fh_matrix* M, GamgHierarchy* H, int mode)
{
    fh_matrix* C = H->mats.back();
    const fh_mesh* cm = C->mesh;
    const int32_t n = cm->nCells, nF = cm->nFaces;
    hipStream_t s = M->stream;
    H->coarseMode = mode;
    if (mode == COARSEST_ITERATIVE)
    {
        H->csDiagN.alloc(n);
        H->csRD.alloc(n);
        H->csSumA.alloc(n);
        H->csWork.alloc(13*(size_t)n);
        CoarsestSys S{};
        S.n = n;
        S.nF = nF;
        S.lower = cm->d_lower.p;
        S.upper = cm->d_upper.p;
        S.rowStart = H->agg->csRowStart.p;
        S.ent = H->agg->csEnt.p;
        S.colN = H->agg->csCol.p;
        S.perm = cm->d_perm.p;
        S.diagLib = C->d_diag.p;
        S.upperF = C->d_upperF.p;
        S.lowerF = C->symmetric ? C->d_upperF.p : C->d_lowerF.p;
        S.diagN = H->csDiagN.p;
        S.rD = H->csRD.p;
        S.sumA = H->csSumA.p;
        launch_coarsest_setup(S, s);
    }
    else
    {
        // the dense coarsest matrix on the device (LUscalarMatrix.C:66-200
        // convert), library order, column-major: no host copy
        const size_t nn = (size_t)n*n;
        H->dA.alloc(nn);
        FH_HIP(hipMemsetAsync(H->dA.p, 0, sizeof(double)*nn, s));
        launch_dense_fill(n, C->rows(), C->d_diag.p, C->d_val.p, H->dA.p, s);
        if (mode == COARSEST_LU)
        {
            // Crout with implicit pivoting in the reference's order
            // (scalarMatrices.C:77-180), one workgroup
            H->luLU.alloc(nn);
            H->luPivot.alloc(n);
            H->luWork.alloc(n);
            H->dFlag.alloc(1);
            FH_HIP(hipMemsetAsync(H->dFlag.p, 0, sizeof(int32_t), s));
            launch_crout_lu(n, H->dA.p, cm->d_perm.p, H->luLU.p, H->luPivot.p, H->dFlag.p, s);
            int32_t bad = 0;
            FH_HIP(hipMemcpyAsync(&bad, H->dFlag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            FH_HIP(hipStreamSynchronize(s));
            if (bad) fail(FH_ERR_NUMERIC, "singular coarsest-level matrix (a zero row)");
        }
        else
        {
            // explicit inverse by Newton-Schulz iteration X <- X (2I - A X)
            // from X0 = A^T/(|A|_1 |A|_inf): FP64 MFMA products, the
            // convergence test in the device state; the host polls it between
            // batches of iterations
            H->dX.alloc(nn);
            H->dT.alloc(nn);
            H->dX1.alloc(nn);
            H->dPart.alloc((nn + 255)/256 + 1);
            H->dNs.alloc(1);
            H->dGemv.alloc((size_t)std::max<int64_t>(gemv_work_doubles(n), 2*(int64_t)n));
            launch_ns_init(n, H->dA.p, H->dX.p, H->dGemv.p, H->dNs.p, s);
            NsState ns{};
            for (int it = 0; it < 200; it += 8)
            {
                launch_ns_iterations(n, H->dA.p, H->dX.p, H->dX1.p, H->dT.p, H->dPart.p,
                                     H->dNs.p, 8, s);
                FH_HIP(hipMemcpyAsync(&ns, H->dNs.p, sizeof(NsState), hipMemcpyDeviceToHost, s));
                FH_HIP(hipStreamSynchronize(s));
                if (ns.done) break;
            }
            if (ns.bad && !(ns.scale > 0.0)) fail(FH_ERR_NUMERIC, "zero coarsest-level matrix");
            if (ns.bad) fail(FH_ERR_NUMERIC, "coarsest inverse diverged");
            if (!(ns.err < 1e-6))
                fail(FH_ERR_NUMERIC, "coarsest inverse did not converge (|I - AX| = %g)", ns.err);
            launch_ns_final(n, H->dX1.p, H->dX.p, H->dNs.p, s);
        }
    }
    H->coarseValid = true;
}

// the agglomeration cache of a mesh and the process-global pairing
// direction are shared by every matrix (and server thread) on the mesh
static std::mutex g_aggMutex;

void gamg_set_face_weights(fh_mesh* mesh, std::vector<double>&& w)
{
    std::lock_guard<std::mutex> lk(g_aggMutex);
    mesh->faceAreaWeights = std::move(w);
    mesh->gamgAgg.reset();
}

// the mesh's cached agglomeration for these controls, built on first use
static std::shared_ptr<GamgAgglomeration> agglomeration_of(fh_matrix* M, const fh_controls* c)
{
    fh_mesh* mesh = M->mesh;
    std::lock_guard<std::mutex> lk(g_aggMutex);
    auto agg = std::static_pointer_cast<GamgAgglomeration>(mesh->gamgAgg);
    if (!agg || agg->minCells != c->nCellsInCoarsestLevel
        || agg->mergeLevels != c->mergeLevels || agg->kind != c->agglomerator)
    {
        agg = build_agglomeration(M, c);
        mesh->gamgAgg = agg;
    }
    return agg;
}

void gamg_prepare(fh_matrix* M, const fh_controls* c)
{
    std::shared_ptr<GamgAgglomeration> agg = agglomeration_of(M, c);
    GamgHierarchy* H = static_cast<GamgHierarchy*>(M->gamg);
    if (H && H->agg != agg)
    {
        delete H;
        H = nullptr;
        M->gamg = nullptr;
    }
    if (!H)
    {
        H = new GamgHierarchy();
        M->gamg = H;
        H->agg = agg;
        for (auto& L : agg->levels)
        {
            fh_matrix* A = new fh_matrix();
            try { matrix_init(A, L->mesh, M->stream); }
            catch (...) { matrix_free(A); delete A; throw; }
            H->mats.push_back(A);
        }
    }
    if (agg->levels.empty())
        fail(FH_ERR_INVALID, "No coarse levels created, either matrix too small for GAMG or "
             "nCellsInCoarsestLevel too large. Either choose another solver of reduce "
             "nCellsInCoarsestLevel.");
    H->ctl = *c;
    H->scale = c->scaleCorrection < 0 ? M->symmetric : c->scaleCorrection != 0;
    hipStream_t s = M->stream;
    if (!H->coeffsValid)
    {
        // GAMGSolverAgglomerateMatrix.C:33-193 level by level
        for (size_t l = 0; l < H->mats.size(); l++)
        {
            fh_matrix* fine = l == 0 ? M : H->mats[l - 1];
            fh_matrix* C = H->mats[l];
            const AggLevel& D = *agg->levels[l];
            C->symmetric = M->symmetric;
            C->diagonalOnly = false;
            C->d_diag.alloc(D.nCoarse);
            if (D.nCoarseFaces > 0)
            {
                C->d_upperF.alloc(D.nCoarseFaces);
                if (!C->symmetric) C->d_lowerF.alloc(D.nCoarseFaces);
                else C->d_lowerF.release();
            }
            GalerkinMaps G{D.rStart.p, D.rCells.p, D.iStart.p, D.iFaces.p,
                           D.fStart.p, D.fFaces.p, D.fFlip.p};
            launch_galerkin(D.nCoarse, D.nCoarseFaces, G, fine->d_diag.p, fine->d_upperF.p,
                            fine->symmetric ? nullptr : fine->d_lowerF.p, C->d_diag.p,
                            C->d_upperF.p, C->symmetric ? nullptr : C->d_lowerF.p, s);
            matrix_coeffs_updated(C);
        }
        H->coeffsValid = true;
        H->coarseValid = false;
    }
    const int32_t nC = H->mats.back()->mesh->nCells;
    int mode = COARSEST_ITERATIVE;
    if (c->directSolveCoarsest)
    {
        if (nC > kInverseMaxCells)
            fail(FH_ERR_UNSUPPORTED, "directSolveCoarsest: coarsest level of %d cells exceeds "
                 "the dense limit %d; raise the agglomeration or use the iterative "
                 "coarsest solve", nC, kInverseMaxCells);
        mode = nC <= kLuMaxCells ? COARSEST_LU : COARSEST_INVERSE;
    }
    if (!H->coarseValid || H->coarseMode != mode) setup_coarsest(M, H, mode);
}

// ---------------------------------------------------------------------------
// V-cycle (GAMGSolverSolve.C:148-443) on library-numbered level vectors
// ---------------------------------------------------------------------------

template<class T>
static void solve_coarsest(fh_matrix* M, GamgHierarchy* H, T* corr, const T* src,
                           DevState* st)
{
    fh_matrix* C = H->mats.back();
    const fh_mesh* cm = C->mesh;
    const int32_t n = cm->nCells;
    hipStream_t s = M->stream;
    switch (H->coarseMode)
    {
        case COARSEST_LU:
            launch_lu_backsub<T>(n, H->luLU.p, H->luPivot.p, cm->d_perm.p, H->luWork.p, src, corr,
                                 st, s);
            break;
        case COARSEST_INVERSE:
            launch_gemv<T>(n, H->dX.p, src, corr, H->dGemv.p, st, s);
            break;
        default:
        {
            CoarsestSys S{};
            S.n = n;
            S.nF = cm->nFaces;
            S.lower = cm->d_lower.p;
            S.upper = cm->d_upper.p;
            S.rowStart = H->agg->csRowStart.p;
            S.ent = H->agg->csEnt.p;
            S.colN = H->agg->csCol.p;
            S.perm = cm->d_perm.p;
            S.diagLib = C->d_diag.p;
            S.upperF = C->d_upperF.p;
            S.lowerF = C->symmetric ? C->d_upperF.p : C->d_lowerF.p;
            S.diagN = H->csDiagN.p;
            S.rD = H->csRD.p;
            S.sumA = H->csSumA.p;
            S.work = H->csWork.p;
            // GAMGSolver.C:270-322: the parent's tolerance and relTol, the
            // default iteration limits
            S.tol = H->ctl.tolerance;
            S.relTol = H->ctl.relTol;
            S.maxIter = 1000;
            S.minIter = 0;
            S.sym = C->symmetric ? 1 : 0;
            launch_coarsest_krylov<T>(S, src, corr, st, s);
        }
    }
}

template<class T>
static void vcycle(fh_matrix* M, GamgHierarchy* H, T* psi, const T* source, T* Apsi,
                   T* fCorr, const T* fRes, DevState* st)
{
    const fh_controls& c = H->ctl;
    GamgWork<T>& W = work_of<T>(H);
    const GamgAgglomeration& A = *H->agg;
    const int coarsest = (int)H->mats.size() - 1;
    hipStream_t s = M->stream;
    const int32_t N = M->mesh->nCells;

    launch_restrict<T>(A.levels[0]->nCoarse, A.levels[0]->rStart.p, A.levels[0]->rCells.p,
                       fRes, W.src[0]->p, st, s);
    for (int l = 0; l < coarsest; l++)
    {
        fh_matrix* L = H->mats[l];
        const int32_t n = L->mesh->nCells;
        if (c.nPreSweeps)
        {
            launch_zero_gamg<T>(n, W.corr[l]->p, st, s);
            smooth_t<T>(L, c.smoother, W.corr[l]->p, W.src[l]->p,
                        std::min(c.nPreSweeps + c.preSweepsLevelMultiplier*l, c.maxPreSweeps),
                        st);
            if (H->scale && l < coarsest - 1)
                launch_gamg_scale<T>(L->rows(), diag_for<T>(L), val_for<T>(L, false),
                                     W.corr[l]->p, W.s1.p, W.src[l]->p, L->d_partials.p, st, s);
            launch_amul<T>(L->rows(), diag_for<T>(L), val_for<T>(L, false), W.corr[l]->p,
                           W.s1.p, st, s);
            launch_sub_gamg<T>(n, W.src[l]->p, W.s1.p, st, s);
        }
        const AggLevel& D = *A.levels[l + 1];
        launch_restrict<T>(D.nCoarse, D.rStart.p, D.rCells.p, W.src[l]->p, W.src[l + 1]->p,
                           st, s);
    }

    solve_coarsest<T>(M, H, W.corr[coarsest]->p, W.src[coarsest]->p, st);

    for (int l = coarsest - 1; l >= 0; l--)
    {
        fh_matrix* L = H->mats[l];
        const int32_t n = L->mesh->nCells;
        if (c.nPreSweeps) launch_copy_gamg<T>(n, W.corr[l]->p, W.s2.p, st, s);
        launch_prolong<T>(n, A.levels[l + 1]->prolong.p, W.corr[l + 1]->p, W.corr[l]->p, st, s);
        if (H->scale && l < coarsest - 1)
            launch_gamg_scale<T>(L->rows(), diag_for<T>(L), val_for<T>(L, false), W.corr[l]->p,
                                 W.s1.p, W.src[l]->p, L->d_partials.p, st, s);
        if (c.nPreSweeps) launch_axpy_gamg<T>(n, W.corr[l]->p, W.s2.p, st, s);
        smooth_t<T>(L, c.smoother, W.corr[l]->p, W.src[l]->p,
                    std::min(c.nPostSweeps + c.postSweepsLevelMultiplier*l, c.maxPostSweeps), st);
    }

    launch_prolong<T>(N, A.levels[0]->prolong.p, W.corr[0]->p, fCorr, st, s);
    if (H->scale)
        launch_gamg_scale<T>(M->rows(), diag_for<T>(M), val_for<T>(M, false), fCorr, Apsi, fRes,
                             M->d_partials.p, st, s);
    launch_axpy_gamg<T>(N, psi, fCorr, st, s);
    smooth_t<T>(M, c.smoother, psi, source, c.nFinestSweeps, st);
}

// GAMGPreconditioner.C:87-148
template<class T>
void gamg_apply(fh_matrix* M, const T* r, T* w, DevState* st)
{
    GamgHierarchy* H = static_cast<GamgHierarchy*>(M->gamg);
    if (!H || !H->coeffsValid || !H->coarseValid)
        fail(FH_ERR_INVALID, "GAMG preconditioner applied before its hierarchy was prepared");
    GamgWork<T>& W = ensure_work<T>(M, H);
    const int32_t N = M->mesh->nCells;
    hipStream_t s = M->stream;
    launch_zero_gamg<T>(N, w, st, s);
    launch_copy_gamg<T>(N, r, W.fRes.p, st, s);
    const int32_t nV = std::max(H->ctl.nVcycles, 1);
    for (int32_t cycle = 0; cycle < nV; cycle++)
    {
        vcycle<T>(M, H, w, r, W.Apsi.p, W.fCorr.p, W.fRes.p, st);
        if (cycle < nV - 1)
        {
            launch_amul<T>(M->rows(), diag_for<T>(M), val_for<T>(M, false), w, W.Apsi.p, st, s);
            launch_diff_gamg<T>(N, r, W.Apsi.p, W.fRes.p, st, s);
        }
    }
}

template void gamg_apply<double>(fh_matrix*, const double*, double*, DevState*);
template void gamg_apply<float>(fh_matrix*, const float*, float*, DevState*);

// GAMGSolverSolve.C:31-145.  The mixed form runs each V-cycle in FP32 on
// the residual equation A e = r (e = 0 first, r scaled by its mean
// magnitude), adds e to the FP64 solution and recomputes the FP64 residual
// for the convergence test.
void gamg_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* src,
                fh_perf* perf, bool mixed)
{
    gamg_prepare(M, c);
    GamgHierarchy* H = static_cast<GamgHierarchy*>(M->gamg);
    const int32_t N = M->mesh->nCells;
    DevState* st = M->d_state.p;
    double* part = M->d_partials.p;
    hipStream_t s = M->stream;
    GamgWork<double>& W = ensure_work<double>(M, H);
    state_upload(M, c, c->tolerance, c->relTol, c->maxIter, c->minIter);
    init_fp64(M, psi, src, W.fRes.p, W.Apsi.p);
    if (!mixed)
    {
        run_batches(M, c->maxIter, [&](int32_t)
        {
            vcycle<double>(M, H, psi, src, W.Apsi.p, W.fCorr.p, W.fRes.p, st);
            launch_amul<double>(M->rows(), M->d_diag.p, M->d_val.p, psi, W.Apsi.p, st, s);
            launch_sub_sum_abs(N, W.Apsi.p, src, W.fRes.p, part, st, s);
            launch_finalize_smooth(N, part, st, 1, s);
        });
    }
    else
    {
        GamgWork<float>& X = ensure_work<float>(M, H);
        run_batches(M, c->maxIter, [&](int32_t)
        {
            launch_to_f32_gamg(N, W.fRes.p, X.r.p, st, s);
            launch_zero_gamg<float>(N, X.e.p, st, s);
            vcycle<float>(M, H, X.e.p, X.r.p, X.Apsi.p, X.fCorr.p, X.r.p, st);
            launch_add_f32_gamg(N, psi, X.e.p, st, s);
            launch_amul<double>(M->rows(), M->d_diag.p, M->d_val.p, psi, W.Apsi.p, st, s);
            launch_sub_sum_abs(N, W.Apsi.p, src, W.fRes.p, part, st, s);
            launch_finalize_smooth(N, part, st, 1, s);
        });
    }
    fill_perf(M, perf, mixed ? "GAMG(mixed)" : "GAMG");
}

} // namespace fh

// GAMGAgglomeration's level maps (GAMGAgglomeration.H:325-367) as seen by
// the C-ABI
namespace
{
struct AggQuery
{
    fh_matrix* A;
    const fh_controls* c;
    int32_t level;
    int32_t* out;
    int32_t *restrictMap, *faceRestrict, *lower, *upper;
    uint8_t* faceFlip;
};
} // namespace

extern "C" int fh_gamg_agglomeration_levels(fh_matrix* A, const fh_controls* c,
                                            int32_t* nLevels)
{
    AggQuery q{A, c, 0, nLevels, nullptr, nullptr, nullptr, nullptr, nullptr};
    return fh::guarded_call([](void* p) {
        AggQuery& q = *static_cast<AggQuery*>(p);
        if (!q.A || !q.c || !q.out) fh::fail(FH_ERR_INVALID, "NULL argument");
        if (q.A->mesh->nInterfaces > 0)
            fh::fail(FH_ERR_UNSUPPORTED, "GAMG agglomeration of coupled interfaces");
        *q.out = (int32_t)fh::agglomeration_of(q.A, q.c)->levels.size();
    }, &q);
}

extern "C" int fh_gamg_agglomeration_level(fh_matrix* A, const fh_controls* c, int32_t level,
                                           int32_t* sizes, int32_t* restrictMap,
                                           int32_t* faceRestrict, uint8_t* faceFlip,
                                           int32_t* lower, int32_t* upper)
{
    AggQuery q{A, c, level, sizes, restrictMap, faceRestrict, lower, upper, faceFlip};
    return fh::guarded_call([](void* p) {
        AggQuery& q = *static_cast<AggQuery*>(p);
        if (!q.A || !q.c || !q.out) fh::fail(FH_ERR_INVALID, "NULL argument");
        if (q.A->mesh->nInterfaces > 0)
            fh::fail(FH_ERR_UNSUPPORTED, "GAMG agglomeration of coupled interfaces");
        auto agg = fh::agglomeration_of(q.A, q.c);
        if (q.level < 0 || q.level >= (int32_t)agg->levels.size())
            fh::fail(FH_ERR_INVALID, "agglomeration level %d of %zu", q.level,
                     agg->levels.size());
        const fh::AggMaps& m = *agg->levels[q.level]->maps;
        q.out[0] = m.nFine;
        q.out[1] = m.nFineFaces;
        q.out[2] = m.nCoarse;
        q.out[3] = m.nCoarseFaces;
        hipStream_t s = q.A->stream;
        auto get = [&](auto* dst, const auto& src, size_t n) {
            if (dst && n)
                FH_HIP(hipMemcpyAsync(dst, src.p, n*sizeof(*dst), hipMemcpyDeviceToHost, s));
        };
        get(q.restrictMap, m.restrictMap, (size_t)m.nFine);
        get(q.faceRestrict, m.faceRestrict, (size_t)m.nFineFaces);
        get(q.faceFlip, m.faceFlip, (size_t)m.nFineFaces);
        get(q.lower, m.lower, (size_t)m.nCoarseFaces);
        get(q.upper, m.upper, (size_t)m.nCoarseFaces);
        FH_HIP(hipStreamSynchronize(s));
    }, &q);
}

// pairGAMGAgglomeration::forward_ as seen by the C-ABI
extern "C" int fh_gamg_direction(void)
{
    return fh::gamg_forward() ? 1 : 0;
}

extern "C" void fh_gamg_reset_direction(void)
{
    fh::gamg_set_forward(true);
}
