// C-ABI of the device assembly kernels (fh_assembly.hip).
#include "fh_assembly.h"
#include "fh_matrix.h"
#include "fh_hops.h"

#include <vector>

struct fh_boundary
{
    fh_mesh* mesh = nullptr;
    int32_t nB = 0;
    fh::DBuf<int32_t> d_faceCells, d_cellStart, d_cellFaces;

    fh::BoundaryView view() const
    {
        fh::BoundaryView v;
        v.nB = nB;
        v.faceCells = d_faceCells.p;
        v.cellStart = d_cellStart.p;
        v.cellFaces = d_cellFaces.p;
        return v;
    }
};

namespace fh
{
int guarded_call(void (*fn)(void*), void* ctx);
void matrix_coeffs_updated(fh_matrix* M);
}

using namespace fh;

namespace
{
template<class F>
int run(F&& f)
{
    return guarded_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

void check_boundary(const fh_boundary* b, const fh_mesh* mesh)
{
    if (!b) fail(FH_ERR_INVALID, "NULL boundary handle");
    if (b->mesh != mesh) fail(FH_ERR_INVALID, "boundary belongs to a different mesh");
}
} // namespace

extern "C" {

int fh_boundary_create(fh_mesh* mesh, int32_t nB, const int32_t* bFaceCells, fh_boundary** out)
{
    return run([&] {
        if (!mesh || !out) fail(FH_ERR_INVALID, "NULL argument");
        if (nB < 0 || (nB > 0 && !bFaceCells)) fail(FH_ERR_INVALID, "bad boundary face list");
        *out = nullptr;
        const int32_t N = mesh->nCells;
        std::vector<int32_t> start(N + 1, 0), faces(nB);
        for (int32_t b = 0; b < nB; b++)
        {
            const int32_t c = bFaceCells[b];
            if (c < 0 || c >= N) fail(FH_ERR_INVALID, "boundary face %d: cell %d out of range", b, c);
            start[c + 1]++;
        }
        for (int32_t c = 0; c < N; c++) start[c + 1] += start[c];
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int32_t b = 0; b < nB; b++) faces[fill[bFaceCells[b]]++] = b;
        fh_boundary* B = new fh_boundary();
        B->mesh = mesh;
        B->nB = nB;
        try
        {
            B->d_faceCells.upload(bFaceCells, nB, mesh->setupStream);
            B->d_cellStart.upload(start, mesh->setupStream);
            B->d_cellFaces.upload(faces, mesh->setupStream);
            FH_HIP(hipStreamSynchronize(mesh->setupStream));
        }
        catch (...) { delete B; throw; }
        *out = B;
    });
}

int fh_boundary_destroy(fh_boundary* b)
{
    return run([&] { delete b; });
}

int fh_assemble_laplacian(fh_matrix* M, const fh_boundary* b, const double* d_gamma,
                          const double* d_weights, const double* d_magSf,
                          const double* d_deltaCoeffs, const double* d_bMagSf,
                          const double* d_bDeltaCoeffs, const int32_t* d_bType,
                          const double* d_bValue, double* d_ic, double* d_bc,
                          double* d_source, int32_t addBoundaryDiag)
{
    return run([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        check_boundary(b, M->mesh);
        fh_mesh* mesh = M->mesh;
        if (!d_gamma || (mesh->nFaces && (!d_weights || !d_magSf || !d_deltaCoeffs)))
            fail(FH_ERR_INVALID, "NULL face/cell input");
        if (b->nB && (!d_bMagSf || !d_bDeltaCoeffs || !d_bType || !d_bValue || !d_ic || !d_bc))
            fail(FH_ERR_INVALID, "NULL boundary input/output");
        LaplacianInputs in{d_gamma, d_weights, d_magSf, d_deltaCoeffs, d_bMagSf,
                           d_bDeltaCoeffs, d_bValue, d_bType};
        M->d_upperF.alloc(std::max(mesh->nFaces, 1));
        M->d_lowerF.release();
        M->d_diag.alloc(mesh->nCells);
        M->symmetric = true;
        M->diagonalOnly = false;
        launch_laplacian(M->rows(), mesh->d_perm.p, mesh->d_lower.p, mesh->d_upper.p,
                         mesh->nFaces, b->view(), in, M->d_upperF.p, M->d_diag.p, d_ic, d_bc,
                         d_source, addBoundaryDiag, M->stream);
        matrix_coeffs_updated(M);
    });
}

int fh_assemble_convection_diffusion(fh_matrix* M, const fh_boundary* b, const double* d_phi,
                                     double D, const double* d_magSf,
                                     const double* d_deltaCoeffs, const double* d_bPhi,
                                     const double* d_bMagSf, const double* d_bDeltaCoeffs,
                                     const int32_t* d_bType, const double* d_bValue,
                                     double* d_ic, double* d_bc, double* d_source,
                                     int32_t addBoundaryDiag)
{
    return run([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        check_boundary(b, M->mesh);
        fh_mesh* mesh = M->mesh;
        if (mesh->nFaces && (!d_phi || !d_magSf || !d_deltaCoeffs))
            fail(FH_ERR_INVALID, "NULL face input");
        if (b->nB && (!d_bPhi || !d_bMagSf || !d_bDeltaCoeffs || !d_bType || !d_bValue
                      || !d_ic || !d_bc))
            fail(FH_ERR_INVALID, "NULL boundary input/output");
        ConvDiffInputs in{d_phi, D, d_magSf, d_deltaCoeffs, d_bPhi, d_bMagSf, d_bDeltaCoeffs,
                          d_bValue, d_bType};
        const int32_t F = std::max(mesh->nFaces, 1);
        M->d_upperF.alloc(F);
        M->d_lowerF.alloc(F);
        M->d_diag.alloc(mesh->nCells);
        M->symmetric = false;
        M->diagonalOnly = false;
        DBuf<double> work;
        work.alloc(3*(size_t)F);
        launch_convection_diffusion(M->rows(), mesh->d_perm.p, mesh->nFaces, b->view(), in,
                                    M->d_upperF.p, M->d_lowerF.p, work.p, M->d_diag.p, d_ic,
                                    d_bc, d_source, addBoundaryDiag, M->stream);
        matrix_coeffs_updated(M);
    });
}

int fh_relax(fh_matrix* M, const fh_boundary* b, double alpha, int32_t nCmpt,
             const double* d_ic, const double* d_psi, double* d_source, double* d_diag)
{
    return run([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        check_boundary(b, M->mesh);
        if (!M->hasCoeffs) fail(FH_ERR_INVALID, "matrix has no coefficients");
        if (nCmpt < 1 || nCmpt > 9) fail(FH_ERR_INVALID, "nCmpt must be in 1..9, got %d", nCmpt);
        if (alpha <= 0) return;
        if (!d_psi || !d_source || (b->nB && !d_ic)) fail(FH_ERR_INVALID, "NULL argument");
        launch_relax(M->rows(), M->mesh->d_perm.p, M->d_val.p, b->view(), d_ic, nCmpt, alpha,
                     d_psi, M->d_diag.p, d_source, d_diag, M->stream);
        // the diagonal changed: drop the cached factors
        matrix_coeffs_updated(M);
    });
}

int fh_flux(fh_matrix* M, const fh_boundary* b, int32_t nCmpt, const double* d_psi,
            const double* d_ic, const double* d_bc, const double* d_corr, const double* d_bCorr,
            double* d_flux, double* d_bFlux)
{
    return run([&] {
        if (!M) fail(FH_ERR_INVALID, "NULL matrix handle");
        check_boundary(b, M->mesh);
        if (!M->hasCoeffs) fail(FH_ERR_INVALID, "matrix has no coefficients");
        if (nCmpt < 1 || nCmpt > 9) fail(FH_ERR_INVALID, "nCmpt must be in 1..9, got %d", nCmpt);
        if (M->diagonalOnly)
            fail(FH_ERR_INVALID, "Cannot calculate faceH the matrix does not have any "
                                 "off-diagonal coefficients.");
        fh_mesh* mesh = M->mesh;
        if (!d_psi || (mesh->nFaces && !d_flux)
            || (b->nB && (!d_ic || !d_bc || !d_bFlux)))
            fail(FH_ERR_INVALID, "NULL argument");
        const double* lower = M->symmetric ? M->d_upperF.p : M->d_lowerF.p;
        launch_flux(mesh->nFaces, mesh->d_lower.p, mesh->d_upper.p, M->d_upperF.p, lower,
                    b->view(), nCmpt, d_psi, d_ic, d_bc, d_corr, d_bCorr, d_flux, d_bFlux,
                    M->stream);
        FH_HIP(hipStreamSynchronize(M->stream));
    });
}

int fh_grad_gauss_linear(fh_mesh* mesh, const fh_boundary* b, const double* d_phi,
                         const double* d_weights, const double* d_Sf, const double* d_V,
                         const double* d_bSf, const double* d_bMagSf, const double* d_bPhi,
                         const double* d_bSnGrad, double* d_grad, double* d_bGrad)
{
    return run([&] {
        if (!mesh) fail(FH_ERR_INVALID, "NULL mesh handle");
        check_boundary(b, mesh);
        if (!d_phi || !d_V || !d_grad || (mesh->nFaces && (!d_weights || !d_Sf)))
            fail(FH_ERR_INVALID, "NULL input/output");
        if (b->nB && (!d_bSf || !d_bPhi)) fail(FH_ERR_INVALID, "NULL boundary input");
        if (d_bGrad && b->nB && (!d_bMagSf || !d_bSnGrad))
            fail(FH_ERR_INVALID, "NULL boundary input for the patch gradient");
        GradInputs in{d_phi, d_weights, d_Sf, d_V, d_bSf, d_bMagSf, d_bPhi, d_bSnGrad};
        DBuf<double> work;
        work.alloc(3*(size_t)std::max(mesh->nFaces, 1));
        RowsView R;
        R.n = mesh->nCells;
        R.rowInfo = mesh->d_rowInfo.p;
        R.sliceStart = mesh->d_sliceStart.p;
        R.col = mesh->d_col.p;
        R.face = mesh->d_face.p;
        launch_grad(R, mesh->d_perm.p, mesh->d_lower.p, mesh->d_upper.p, mesh->nFaces,
                    b->view(), in, work.p, d_grad, d_bGrad, mesh->setupStream);
        FH_HIP(hipStreamSynchronize(mesh->setupStream));
    });
}

int fh_face_geometry(int32_t nFaces, const int32_t* d_faceStart, const int32_t* d_facePoints,
                     const double* d_points, double* d_Cf, double* d_Sf, double* d_magSf)
{
    return run([&] {
        if (nFaces < 0) fail(FH_ERR_INVALID, "nFaces must be >= 0");
        if (nFaces > 0 && (!d_faceStart || !d_facePoints || !d_points || !d_Cf || !d_Sf || !d_magSf))
            fail(FH_ERR_INVALID, "NULL input/output");
        launch_face_geometry(nFaces, d_faceStart, d_facePoints, d_points, d_Cf, d_Sf, d_magSf,
                             nullptr);
        FH_HIP(hipStreamSynchronize(nullptr));
    });
}

int fh_cell_geometry(fh_mesh* mesh, const fh_boundary* b, const double* d_Cf, const double* d_Sf,
                     const double* d_bCf, const double* d_bSf, double* d_C, double* d_V)
{
    return run([&] {
        if (!mesh) fail(FH_ERR_INVALID, "NULL mesh handle");
        check_boundary(b, mesh);
        if (!d_C || !d_V || (mesh->nFaces && (!d_Cf || !d_Sf)))
            fail(FH_ERR_INVALID, "NULL input/output");
        if (b->nB && (!d_bCf || !d_bSf)) fail(FH_ERR_INVALID, "NULL boundary input");
        RowsView R;
        R.n = mesh->nCells;
        R.rowInfo = mesh->d_rowInfo.p;
        R.sliceStart = mesh->d_sliceStart.p;
        R.col = mesh->d_col.p;
        R.face = mesh->d_face.p;
        launch_cell_geometry(R, mesh->d_perm.p, b->view(), d_Cf, d_Sf, d_bCf, d_bSf, d_C, d_V,
                             mesh->setupStream);
        FH_HIP(hipStreamSynchronize(mesh->setupStream));
    });
}

int fh_surface_interpolation(fh_mesh* mesh, const fh_boundary* b, const double* d_C,
                             const double* d_Cf, const double* d_Sf, const double* d_magSf,
                             const double* d_bCf, const double* d_bSf, const double* d_bMagSf,
                             double* d_weights, double* d_deltaCoeffs,
                             double* d_nonOrthDeltaCoeffs, double* d_bDeltaCoeffs,
                             double* d_bNonOrthDeltaCoeffs)
{
    return run([&] {
        if (!mesh) fail(FH_ERR_INVALID, "NULL mesh handle");
        check_boundary(b, mesh);
        if (!d_C || (mesh->nFaces && (!d_Cf || !d_Sf || !d_magSf || !d_weights || !d_deltaCoeffs
                                      || !d_nonOrthDeltaCoeffs)))
            fail(FH_ERR_INVALID, "NULL input/output");
        if (b->nB && (!d_bCf || !d_bSf || !d_bMagSf || !d_bDeltaCoeffs))
            fail(FH_ERR_INVALID, "NULL boundary input/output");
        launch_surface_interpolation(mesh->nFaces, mesh->d_lower.p, mesh->d_upper.p, d_C, d_Cf,
                                     d_Sf, d_magSf, d_weights, d_deltaCoeffs,
                                     d_nonOrthDeltaCoeffs, b->nB, b->d_faceCells.p, d_bCf, d_bSf,
                                     d_bMagSf, d_bDeltaCoeffs, d_bNonOrthDeltaCoeffs,
                                     mesh->setupStream);
        FH_HIP(hipStreamSynchronize(mesh->setupStream));
    });
}

} // extern "C"
