// C-ABI entry points of libfoamhip (declared in include/foamhip.h).  Every
// function converts internal fh::Error exceptions into a status code and
// records the message for fh_last_error().
#include "fh_matrix.h"
#include "fh_gamg.h"
#include "fh_hops.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>
#include <strings.h>

namespace fh
{

static thread_local std::string g_lastError;
static int g_device = 0;

void set_last_error(const std::string& msg) { g_lastError = msg; }

void fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error{code, std::string(buf)};
}

template<class F>
static int guarded(F&& f)
{
    try
    {
        f();
        return FH_OK;
    }
    catch (const Error& e)
    {
        set_last_error(e.msg);
        return e.code;
    }
    catch (const std::bad_alloc&)
    {
        set_last_error("host allocation failed");
        return FH_ERR_NOMEM;
    }
    catch (const std::exception& e)
    {
        set_last_error(e.what());
        return FH_ERR_INVALID;
    }
}

int guarded_call(void (*fn)(void*), void* ctx)
{
    return guarded([&] { fn(ctx); });
}

struct NameId
{
    const char* name;
    int32_t id;
};

static int lookup(const NameId* table, const char* name, int32_t* id, const char* what)
{
    if (!name || !id) fail(FH_ERR_INVALID, "NULL argument to %s lookup", what);
    const char* n = name;
    if (std::strncmp(n, "hip", 3) == 0) n += 3;
    for (const NameId* t = table; t->name; t++)
    {
        if (std::strcmp(t->name, n) == 0)
        {
            *id = t->id;
            return FH_OK;
        }
    }
    std::string valid;
    for (const NameId* t = table; t->name; t++) { valid += ' '; valid += t->name; }
    fail(FH_ERR_INVALID, "Unknown %s %s; valid %ss:%s", what, name, what, valid.c_str());
}

// names of the reference run-time selection tables
static const NameId kSolvers[] = {
    {"PCG", FH_SOLVER_PCG}, {"PBiCGStab", FH_SOLVER_PBICGSTAB}, {"PBiCG", FH_SOLVER_PBICG},
    {"smoothSolver", FH_SOLVER_SMOOTH}, {"GAMG", FH_SOLVER_GAMG},
    {"diagonal", FH_SOLVER_DIAGONAL}, {nullptr, 0}};
static const NameId kPrecons[] = {
    {"none", FH_PRECON_NONE}, {"diagonal", FH_PRECON_DIAGONAL}, {"DIC", FH_PRECON_DIC},
    {"FDIC", FH_PRECON_FDIC}, {"DILU", FH_PRECON_DILU}, {"GAMG", FH_PRECON_GAMG},
    {"tileDIC", FH_PRECON_TILE_DIC}, {"tileDILU", FH_PRECON_TILE_DILU}, {nullptr, 0}};
static const NameId kSmoothers[] = {
    {"GaussSeidel", FH_SMOOTH_GS}, {"nonBlockingGaussSeidel", FH_SMOOTH_GS},
    {"symGaussSeidel", FH_SMOOTH_SYMGS}, {"DIC", FH_SMOOTH_DIC}, {"DILU", FH_SMOOTH_DILU},
    {"DICGaussSeidel", FH_SMOOTH_DICGS}, {"DILUGaussSeidel", FH_SMOOTH_DILUGS},
    {"FDIC", FH_SMOOTH_FDIC}, {"none", FH_SMOOTH_NONE}, {nullptr, 0}};

static void check_matrix(const fh_matrix* m)
{
    if (!m) fail(FH_ERR_INVALID, "NULL matrix handle");
    if (!m->hasCoeffs) fail(FH_ERR_INVALID, "matrix coefficients not set");
}

// host-array helpers: upload n doubles in natural order into a library-
// numbered device vector, and back
static void to_lib(fh_matrix* M, const double* h, DBuf<double>& nat, double* lib)
{
    const int32_t N = M->mesh->nCells;
    nat.alloc(N);
    FH_HIP(hipMemcpyAsync(nat.p, h, sizeof(double)*N, hipMemcpyHostToDevice, M->stream));
    launch_permute_in<double, double>(N, M->mesh->d_perm.p, nat.p, lib, M->stream);
}

static void from_lib(fh_matrix* M, const double* lib, DBuf<double>& nat, double* h)
{
    const int32_t N = M->mesh->nCells;
    nat.alloc(N);
    launch_permute_out<double, double>(N, M->mesh->d_perm.p, lib, nat.p, M->stream);
    FH_HIP(hipMemcpyAsync(h, nat.p, sizeof(double)*N, hipMemcpyDeviceToHost, M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
}

} // namespace fh

using namespace fh;

extern "C" {

int fh_version(void) { return 10000; }

const char* fh_last_error(void) { return g_lastError.c_str(); }

int fh_set_device(int32_t device)
{
    return guarded([&] {
        FH_HIP(hipSetDevice(device));
        g_device = device;
    });
}

void fh_controls_default(fh_controls* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->solver = FH_SOLVER_PCG;
    c->preconditioner = FH_PRECON_DIC;
    c->smoother = FH_SMOOTH_GS;
    c->precision = FH_PRECISION_DOUBLE;
    c->maxIter = 1000;                  // lduMatrix::solver::defaultMaxIter_
    c->minIter = 0;
    c->tolerance = 1e-6;
    c->relTol = 0.0;
    c->nSweeps = 1;
    c->nPreSweeps = 0;
    c->preSweepsLevelMultiplier = 1;
    c->maxPreSweeps = 4;
    c->nPostSweeps = 2;
    c->postSweepsLevelMultiplier = 1;
    c->maxPostSweeps = 4;
    c->nFinestSweeps = 2;
    c->scaleCorrection = -1;
    c->directSolveCoarsest = 0;
    c->nCellsInCoarsestLevel = 10;
    c->mergeLevels = 1;
    c->nVcycles = 2;
    c->innerRelTol = 1e-4;
    c->maxRefinements = 50;
    c->innerPreconditioner = -1;
    c->agglomerator = FH_AGGLOM_ALGEBRAIC_PAIR;
    c->preconditionerControls = nullptr;
}

int fh_solver_lookup(const char* name, int32_t* id)
{
    return guarded([&] { lookup(kSolvers, name, id, "solver"); });
}

int fh_preconditioner_lookup(const char* name, int32_t* id)
{
    return guarded([&] { lookup(kPrecons, name, id, "preconditioner"); });
}

int fh_smoother_lookup(const char* name, int32_t* id)
{
    return guarded([&] { lookup(kSmoothers, name, id, "smoother"); });
}

int fh_mesh_create(int32_t nCells, int32_t nFaces, const int32_t* lowerAddr,
                   const int32_t* upperAddr, fh_mesh** out)
{
    return guarded([&] {
        if (!out) fail(FH_ERR_INVALID, "NULL output handle");
        *out = nullptr;
        fh_mesh* m = new fh_mesh();
        try
        {
            m->device = g_device;
            FH_HIP(hipSetDevice(m->device));
            FH_HIP(hipStreamCreateWithFlags(&m->setupStream, hipStreamNonBlocking));
            build_mesh(m, nCells, nFaces, lowerAddr, upperAddr);
        }
        catch (...)
        {
            if (m->setupStream) (void)hipStreamDestroy(m->setupStream);
            delete m;
            throw;
        }
        *out = m;
    });
}

int fh_mesh_create_coupled(int32_t nCells, int32_t nFaces, const int32_t* lowerAddr,
                           const int32_t* upperAddr, int32_t nInterfaces,
                           const int32_t* interfaceStart, const int32_t* faceCells,
                           const int32_t* nbrFaceCells, fh_mesh** out)
{
    return guarded([&] {
        if (!out) fail(FH_ERR_INVALID, "NULL output handle");
        *out = nullptr;
        fh_mesh* m = new fh_mesh();
        try
        {
            m->device = g_device;
            FH_HIP(hipSetDevice(m->device));
            FH_HIP(hipStreamCreateWithFlags(&m->setupStream, hipStreamNonBlocking));
            build_mesh(m, nCells, nFaces, lowerAddr, upperAddr, nInterfaces == 0);
            build_interfaces(m, nInterfaces, interfaceStart, faceCells, nbrFaceCells);
        }
        catch (...)
        {
            if (m->setupStream) (void)hipStreamDestroy(m->setupStream);
            delete m;
            throw;
        }
        *out = m;
    });
}

int fh_mesh_set_face_areas(fh_mesh* mesh, const double* Sf, const double* magSf)
{
    return guarded([&] {
        if (!mesh) fail(FH_ERR_INVALID, "NULL mesh handle");
        const int32_t F = mesh->nFaces;
        if (F > 0 && (!Sf || !magSf)) fail(FH_ERR_INVALID, "NULL face areas");
        // mag(cmptMultiply(Sf/sqrt(magSf), (1, 1.01, 1.02))), the reference's
        // operation order: each component divided by sqrt(magSf), scaled,
        // then sqrt((x*x + y*y) + z*z)
        std::vector<double> w(F);
        for (int32_t f = 0; f < F; f++)
        {
            const double s = std::sqrt(magSf[f]);
            const double x = (Sf[3*f]/s)*1.0, y = (Sf[3*f + 1]/s)*1.01,
                         z = (Sf[3*f + 2]/s)*1.02;
            w[f] = std::sqrt(x*x + y*y + z*z);
        }
        gamg_set_face_weights(mesh, std::move(w));
    });
}

int fh_mesh_info_get(const fh_mesh* m, fh_mesh_info* info)
{
    return guarded([&] {
        if (!m || !info) fail(FH_ERR_INVALID, "NULL argument");
        info->nCells = m->nCells;
        info->nFaces = m->nFaces;
        info->nLevels = m->sched.nLevels;
        info->nParts = m->sched.nParts;
        info->boxNx = m->boxNx;
        info->boxNy = m->boxNy;
        info->boxNz = m->boxNz;
        info->tiled = m->tiled ? 1 : 0;
        info->deviceBytes = (int64_t)mesh_device_bytes(m);
    });
}

int fh_mesh_destroy(fh_mesh* m)
{
    return guarded([&] {
        if (!m) return;
        if (m->setupStream) (void)hipStreamDestroy(m->setupStream);
        delete m;
    });
}

int fh_matrix_create(fh_mesh* mesh, fh_matrix** out)
{
    return guarded([&] {
        if (!mesh || !out) fail(FH_ERR_INVALID, "NULL argument");
        *out = nullptr;
        fh_matrix* M = new fh_matrix();
        try { matrix_init(M, mesh); }
        catch (...) { matrix_free(M); delete M; throw; }
        *out = M;
    });
}

int fh_matrix_set(fh_matrix* M, const double* diag, const double* upper, const double* lower)
{
    return guarded([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        matrix_set(M, diag, upper, lower, false);
    });
}

int fh_matrix_set_device(fh_matrix* M, const double* d_diag, const double* d_upper,
                         const double* d_lower)
{
    return guarded([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        matrix_set(M, d_diag, d_upper, d_lower, true);
    });
}

int fh_matrix_set_interfaces(fh_matrix* M, const double* interfaceBouCoeffs,
                             const double* interfaceIntCoeffs)
{
    return guarded([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        const int32_t nE = M->mesh->nIfFaces;
        if (nE == 0) fail(FH_ERR_INVALID, "the mesh has no coupled interfaces");
        if (!interfaceBouCoeffs) fail(FH_ERR_INVALID, "NULL interfaceBouCoeffs");
        FH_HIP(hipSetDevice(M->mesh->device));
        M->d_ifBou.upload(interfaceBouCoeffs, nE, M->stream);
        M->d_ifInt.upload(interfaceIntCoeffs ? interfaceIntCoeffs : interfaceBouCoeffs, nE,
                          M->stream);
        M->ifSet = true;
        if (M->hasCoeffs) matrix_coeffs_updated(M);
        else FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_matrix_get(fh_matrix* M, double* diag, double* upper, double* lower)
{
    return guarded([&] {
        check_matrix(M);
        const int32_t F = M->mesh->nFaces;
        if (diag)
        {
            DBuf<double> nat;
            from_lib(M, M->d_diag.p, nat, diag);
        }
        auto face = [&](double* h, const double* d) {
            if (!h) return;
            if (M->diagonalOnly || F == 0) std::fill(h, h + F, 0.0);
            else FH_HIP(hipMemcpyAsync(h, d, sizeof(double)*F, hipMemcpyDeviceToHost, M->stream));
        };
        face(upper, M->d_upperF.p);
        face(lower, M->symmetric ? M->d_upperF.p : M->d_lowerF.p);
        FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_matrix_destroy(fh_matrix* M)
{
    return guarded([&] {
        if (!M) return;
        matrix_free(M);
        delete M;
    });
}

int fh_solve(fh_matrix* M, const fh_controls* c, double* psi, const double* source,
             fh_perf* perf)
{
    return guarded([&] {
        check_matrix(M);
        if (!c || !psi || !source || !perf) fail(FH_ERR_INVALID, "NULL argument");
        const int32_t N = M->mesh->nCells;
        M->d_psiNat.alloc(N);
        M->d_srcNat.alloc(N);
        FH_HIP(hipMemcpyAsync(M->d_psiNat.p, psi, sizeof(double)*N, hipMemcpyHostToDevice,
                              M->stream));
        FH_HIP(hipMemcpyAsync(M->d_srcNat.p, source, sizeof(double)*N, hipMemcpyHostToDevice,
                              M->stream));
        try
        {
            solve_device(M, c, M->d_psiNat.p, M->d_srcNat.p, perf);
        }
        catch (const Error& e)
        {
            if (e.code != FH_ERR_NUMERIC && e.code != FH_ERR_SCHEDULE) throw;
            FH_HIP(hipMemcpyAsync(psi, M->d_psiNat.p, sizeof(double)*N, hipMemcpyDeviceToHost,
                                  M->stream));
            FH_HIP(hipStreamSynchronize(M->stream));
            throw;
        }
        FH_HIP(hipMemcpyAsync(psi, M->d_psiNat.p, sizeof(double)*N, hipMemcpyDeviceToHost,
                              M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_solve_device(fh_matrix* M, const fh_controls* c, double* d_psi, const double* d_source,
                    fh_perf* perf)
{
    return guarded([&] {
        check_matrix(M);
        if (!c || !d_psi || !d_source || !perf) fail(FH_ERR_INVALID, "NULL argument");
        solve_device(M, c, d_psi, d_source, perf);
    });
}

int fh_amul(fh_matrix* M, const double* psi, double* Apsi)
{
    return guarded([&] {
        check_matrix(M);
        const int32_t N = M->mesh->nCells;
        DBuf<double> nat, x, y;
        x.alloc(N); y.alloc(N);
        to_lib(M, psi, nat, x.p);
        launch_amul<double>(M->rows(), M->d_diag.p, M->d_val.p, x.p, y.p, nullptr, M->stream);
        from_lib(M, y.p, nat, Apsi);
    });
}

int fh_tmul(fh_matrix* M, const double* psi, double* Tpsi)
{
    return guarded([&] {
        check_matrix(M);
        ensure_valT(M);
        const int32_t N = M->mesh->nCells;
        DBuf<double> nat, x, y;
        x.alloc(N); y.alloc(N);
        to_lib(M, psi, nat, x.p);
        launch_tmul<double>(M->rows(), M->d_diag.p, M->d_valT.p, x.p, y.p, M->stream);
        from_lib(M, y.p, nat, Tpsi);
    });
}

int fh_residual(fh_matrix* M, const double* psi, const double* source, double* rA)
{
    return guarded([&] {
        check_matrix(M);
        const int32_t N = M->mesh->nCells;
        DBuf<double> nat, x, b, r;
        x.alloc(N); b.alloc(N); r.alloc(N);
        to_lib(M, psi, nat, x.p);
        to_lib(M, source, nat, b.p);
        launch_residual<double>(M->rows(), M->d_diag.p, M->d_val.p, x.p, b.p, r.p, nullptr,
                                M->stream);
        from_lib(M, r.p, nat, rA);
    });
}

int fh_sum_a(fh_matrix* M, double* sumA)
{
    return guarded([&] {
        check_matrix(M);
        ensure_sumA(M);
        DBuf<double> nat;
        from_lib(M, M->d_sumA.p, nat, sumA);
    });
}

// a field of nCmpt components per cell (component fastest) into and out of
// the library numbering
static void to_lib_cmpt(fh_matrix* M, int32_t nC, const double* h, DBuf<double>& nat,
                        double* lib)
{
    const int32_t N = M->mesh->nCells;
    nat.alloc((size_t)N*nC);
    FH_HIP(hipMemcpyAsync(nat.p, h, sizeof(double)*N*nC, hipMemcpyHostToDevice, M->stream));
    launch_permute_cmpt_in(N, nC, M->mesh->d_perm.p, nat.p, lib, M->stream);
}

static void from_lib_cmpt(fh_matrix* M, int32_t nC, const double* lib, DBuf<double>& nat,
                          double* h)
{
    const int32_t N = M->mesh->nCells;
    nat.alloc((size_t)N*nC);
    launch_permute_cmpt_out(N, nC, M->mesh->d_perm.p, lib, nat.p, M->stream);
    FH_HIP(hipMemcpyAsync(h, nat.p, sizeof(double)*N*nC, hipMemcpyDeviceToHost, M->stream));
    FH_HIP(hipStreamSynchronize(M->stream));
}

int fh_h(fh_matrix* M, int32_t nCmpt, const double* psi, double* Hpsi)
{
    return guarded([&] {
        check_matrix(M);
        if (!psi || !Hpsi) fail(FH_ERR_INVALID, "NULL argument");
        if (nCmpt < 1 || nCmpt > 9) fail(FH_ERR_INVALID, "nCmpt must be in 1..9, got %d", nCmpt);
        const int32_t N = M->mesh->nCells;
        if (M->diagonalOnly || M->mesh->nFaces == 0)
        {
            std::fill(Hpsi, Hpsi + (int64_t)N*nCmpt, 0.0);
            return;
        }
        DBuf<double> nat, x, y;
        x.alloc((size_t)N*nCmpt); y.alloc((size_t)N*nCmpt);
        to_lib_cmpt(M, nCmpt, psi, nat, x.p);
        launch_h(M->rows(), M->d_val.p, nCmpt, x.p, y.p, M->stream);
        from_lib_cmpt(M, nCmpt, y.p, nat, Hpsi);
    });
}

int fh_h1(fh_matrix* M, double* H1)
{
    return guarded([&] {
        check_matrix(M);
        if (!H1) fail(FH_ERR_INVALID, "NULL argument");
        const int32_t N = M->mesh->nCells;
        if (M->diagonalOnly || M->mesh->nFaces == 0)
        {
            std::fill(H1, H1 + N, 0.0);
            return;
        }
        DBuf<double> nat, y;
        y.alloc(N);
        launch_h1(M->rows(), M->d_val.p, y.p, M->stream);
        from_lib(M, y.p, nat, H1);
    });
}

int fh_face_h(fh_matrix* M, int32_t nCmpt, const double* psi, double* faceHpsi)
{
    return guarded([&] {
        check_matrix(M);
        if (!psi || !faceHpsi) fail(FH_ERR_INVALID, "NULL argument");
        if (nCmpt < 1 || nCmpt > 9) fail(FH_ERR_INVALID, "nCmpt must be in 1..9, got %d", nCmpt);
        if (M->diagonalOnly)
            fail(FH_ERR_INVALID, "Cannot calculate faceH the matrix does not have any "
                                 "off-diagonal coefficients.");
        const fh_mesh* mesh = M->mesh;
        const int32_t N = mesh->nCells, F = mesh->nFaces;
        if (F == 0) return;
        DBuf<double> x, y;
        x.alloc((size_t)N*nCmpt); y.alloc((size_t)F*nCmpt);
        FH_HIP(hipMemcpyAsync(x.p, psi, sizeof(double)*N*nCmpt, hipMemcpyHostToDevice,
                              M->stream));
        const double* lower = M->symmetric ? M->d_upperF.p : M->d_lowerF.p;
        launch_face_h(F, mesh->d_lower.p, mesh->d_upper.p, M->d_upperF.p, lower, nCmpt, x.p,
                      y.p, M->stream);
        FH_HIP(hipMemcpyAsync(faceHpsi, y.p, sizeof(double)*F*nCmpt, hipMemcpyDeviceToHost,
                              M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_sum_mag_off_diag(fh_matrix* M, double* sumOff)
{
    return guarded([&] {
        check_matrix(M);
        if (!sumOff) fail(FH_ERR_INVALID, "NULL argument");
        const int32_t N = M->mesh->nCells;
        if (M->diagonalOnly || M->mesh->nFaces == 0) return;
        DBuf<double> nat, y;
        y.alloc(N);
        to_lib(M, sumOff, nat, y.p);
        launch_sum_mag_off_diag(M->rows(), M->d_val.p, y.p, M->stream);
        from_lib(M, y.p, nat, sumOff);
    });
}

int fh_h_device(fh_matrix* M, int32_t nCmpt, const double* d_psi, double* d_Hpsi)
{
    return guarded([&] {
        check_matrix(M);
        if (!d_psi || !d_Hpsi) fail(FH_ERR_INVALID, "NULL argument");
        if (nCmpt < 1 || nCmpt > 9) fail(FH_ERR_INVALID, "nCmpt must be in 1..9, got %d", nCmpt);
        const int32_t N = M->mesh->nCells;
        if (N == 0) return;
        if (M->diagonalOnly || M->mesh->nFaces == 0)
        {
            FH_HIP(hipMemsetAsync(d_Hpsi, 0, sizeof(double)*N*nCmpt, M->stream));
            FH_HIP(hipStreamSynchronize(M->stream));
            return;
        }
        DBuf<double> x, y;
        x.alloc((size_t)N*nCmpt); y.alloc((size_t)N*nCmpt);
        launch_permute_cmpt_in(N, nCmpt, M->mesh->d_perm.p, d_psi, x.p, M->stream);
        launch_h(M->rows(), M->d_val.p, nCmpt, x.p, y.p, M->stream);
        launch_permute_cmpt_out(N, nCmpt, M->mesh->d_perm.p, y.p, d_Hpsi, M->stream);
        FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_face_h_device(fh_matrix* M, int32_t nCmpt, const double* d_psi, double* d_faceHpsi)
{
    return guarded([&] {
        check_matrix(M);
        if (!d_psi || !d_faceHpsi) fail(FH_ERR_INVALID, "NULL argument");
        if (nCmpt < 1 || nCmpt > 9) fail(FH_ERR_INVALID, "nCmpt must be in 1..9, got %d", nCmpt);
        if (M->diagonalOnly)
            fail(FH_ERR_INVALID, "Cannot calculate faceH the matrix does not have any "
                                 "off-diagonal coefficients.");
        const fh_mesh* mesh = M->mesh;
        const double* lower = M->symmetric ? M->d_upperF.p : M->d_lowerF.p;
        launch_face_h(mesh->nFaces, mesh->d_lower.p, mesh->d_upper.p, M->d_upperF.p, lower,
                      nCmpt, d_psi, d_faceHpsi, M->stream);
        FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_norm_factor(fh_matrix* M, const double* psi, const double* source, double* normFactor)
{
    return guarded([&] {
        check_matrix(M);
        if (!normFactor) fail(FH_ERR_INVALID, "NULL argument");
        const int32_t N = M->mesh->nCells;
        DBuf<double> nat, x, b, Ax, r;
        x.alloc(N); b.alloc(N); Ax.alloc(N); r.alloc(N);
        to_lib(M, psi, nat, x.p);
        to_lib(M, source, nat, b.p);
        ensure_sumA(M);
        double* part = M->d_partials.p;
        DevState* st = M->d_state.p;
        std::memset(M->h_state, 0, sizeof(DevState));
        FH_HIP(hipMemcpyAsync(st, M->h_state, sizeof(DevState), hipMemcpyHostToDevice, M->stream));
        launch_amul<double>(M->rows(), M->d_diag.p, M->d_val.p, x.p, Ax.p, nullptr, M->stream);
        launch_sum_abs<double>(N, x.p, part, M->stream);
        launch_finalize(N, part, FIN_XREF, st, 0, M->stream);
        launch_init_residual(N, Ax.p, b.p, M->d_sumA.p, r.p, part, st, M->stream);
        launch_finalize(N, part, FIN_INIT, st, 0, M->stream);
        FH_HIP(hipMemcpyAsync(M->h_state, st, sizeof(DevState), hipMemcpyDeviceToHost, M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
        *normFactor = M->h_state->normFactor;
    });
}

int fh_reciprocal_d(fh_matrix* M, int32_t precon, double* rD)
{
    return guarded([&] {
        check_matrix(M);
        int which = -1;
        if (precon == FH_PRECON_DIC || precon == FH_PRECON_FDIC) which = 0;
        else if (precon == FH_PRECON_DILU) which = 1;
        else if (precon == FH_PRECON_DIAGONAL) which = 2;
        else if (precon == FH_PRECON_TILE_DIC) which = 3;
        else if (precon == FH_PRECON_TILE_DILU) which = 4;
        else fail(FH_ERR_INVALID, "preconditioner %d has no reciprocal factor", precon);
        const double* d = ensure_rD(M, which);
        DBuf<double> nat;
        from_lib(M, d, nat, rD);
        check_sweep_errors(M);
    });
}

static int precondition_impl(fh_matrix* M, int32_t precon, const double* rA, double* wA,
                             bool transpose)
{
    return guarded([&] {
        check_matrix(M);
        if (precon == FH_PRECON_GAMG)
            fail(FH_ERR_UNSUPPORTED, "use fh_solve for the GAMG preconditioner");
        const int32_t N = M->mesh->nCells;
        DBuf<double> nat, r, w;
        r.alloc(N); w.alloc(N);
        to_lib(M, rA, nat, r.p);
        apply_precond<double>(M, precon, r.p, w.p, nullptr, nullptr, transpose);
        from_lib(M, w.p, nat, wA);
        check_sweep_errors(M);
    });
}

int fh_precondition(fh_matrix* M, int32_t precon, const double* rA, double* wA)
{
    return precondition_impl(M, precon, rA, wA, false);
}

int fh_precondition_t(fh_matrix* M, int32_t precon, const double* rA, double* wA)
{
    return precondition_impl(M, precon, rA, wA, true);
}

int fh_precondition_f32(fh_matrix* M, int32_t precon, const float* rA, float* wA)
{
    return guarded([&] {
        check_matrix(M);
        if (!rA || !wA) fail(FH_ERR_INVALID, "NULL argument");
        if (precon == FH_PRECON_GAMG)
            fail(FH_ERR_UNSUPPORTED, "use fh_solve for the GAMG preconditioner");
        const int32_t N = M->mesh->nCells;
        DBuf<float> nat, r, w;
        nat.alloc(N); r.alloc(N); w.alloc(N);
        FH_HIP(hipMemcpyAsync(nat.p, rA, sizeof(float)*N, hipMemcpyHostToDevice, M->stream));
        launch_permute_in<float, float>(N, M->mesh->d_perm.p, nat.p, r.p, M->stream);
        apply_precond<float>(M, precon, r.p, w.p, nullptr, nullptr, false);
        launch_permute_out<float, float>(N, M->mesh->d_perm.p, w.p, nat.p, M->stream);
        FH_HIP(hipMemcpyAsync(wA, nat.p, sizeof(float)*N, hipMemcpyDeviceToHost, M->stream));
        FH_HIP(hipStreamSynchronize(M->stream));
        check_sweep_errors(M);
    });
}

int fh_smooth(fh_matrix* M, int32_t smoother, double* psi, const double* source,
              int32_t nSweeps)
{
    return guarded([&] {
        check_matrix(M);
        const int32_t N = M->mesh->nCells;
        DBuf<double> nat, x, b;
        x.alloc(N); b.alloc(N);
        to_lib(M, psi, nat, x.p);
        to_lib(M, source, nat, b.p);
        smooth_device(M, smoother, x.p, b.p, nSweeps, nullptr);
        from_lib(M, x.p, nat, psi);
        check_sweep_errors(M);
    });
}

int fh_profile_kernel(fh_matrix* M, int32_t kernel, int32_t precision, int32_t precon,
                      int32_t reps, double* ms, double* bytes)
{
    return guarded([&] {
        check_matrix(M);
        if (!ms || !bytes) fail(FH_ERR_INVALID, "NULL output");
        profile_kernel(M, kernel, precision, precon, reps, ms, bytes);
    });
}

int fh_kernel_timing(fh_matrix* M, uint32_t kernelMask)
{
    return guarded([&] {
        check_matrix(M);
        if (kernelMask >> 5) fail(FH_ERR_INVALID, "kernel mask 0x%x has unknown bits", kernelMask);
        auto& T = M->kt;
        T.mask = kernelMask;
        for (int k = 0; k < 5; k++)
        {
            T.used[k] = 0;
            T.ms[k] = 0.0;
            T.launches[k] = 0;
        }
    });
}

int fh_kernel_timing_get(fh_matrix* M, int32_t kernel, double* msTotal, int64_t* launches)
{
    return guarded([&] {
        check_matrix(M);
        if (!msTotal || !launches) fail(FH_ERR_INVALID, "NULL output");
        if (kernel < 0 || kernel > 4) fail(FH_ERR_INVALID, "unknown kernel id %d", kernel);
        *msTotal = M->kt.ms[kernel];
        *launches = M->kt.launches[kernel];
    });
}

int fh_profile_stream(int32_t pattern, int64_t bytes, double* bytesPerSecond)
{
    return guarded([&] {
        if (!bytesPerSecond) fail(FH_ERR_INVALID, "NULL output");
        if (pattern < 0 || pattern > 2) fail(FH_ERR_INVALID, "unknown stream pattern %d", pattern);
        if (bytes < (1 << 20) || (bytes & 15)) fail(FH_ERR_INVALID, "bytes must be >= 1 MiB and a multiple of 16");
        const int nBuf = pattern == 0 ? 1 : (pattern == 1 ? 2 : 4);
        fh::DBuf<float4> buf[4];
        float4* p[4] = {};
        for (int k = 0; k < nBuf; k++)
        {
            buf[k].alloc((size_t)bytes/16);
            FH_HIP(hipMemset(buf[k].p, 0, (size_t)bytes));
            p[k] = buf[k].p;
        }
        fh::DBuf<float> out;
        out.alloc(1);
        hipEvent_t e0, e1;
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        const int64_t n4 = bytes/16;
        fh::launch_stream_probe(pattern, n4, 2048, p, out.p, nullptr);
        for (int rep = 0; rep < 5; rep++)
        {
            FH_HIP(hipEventRecord(e0, nullptr));
            fh::launch_stream_probe(pattern, n4, 2048, p, out.p, nullptr);
            FH_HIP(hipEventRecord(e1, nullptr));
            FH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FH_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double moved = (double)bytes*(pattern == 0 ? 1 : (pattern == 1 ? 2 : 4));
        *bytesPerSecond = moved/(best*1e-3);
    });
}

} // extern "C"
