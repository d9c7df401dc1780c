// Assembly gather kernels (fh_assembly.hip).
#pragma once

#include "fh_kernels.h"

namespace fh
{

// boundary faces grouped per cell in patch/face order
struct BoundaryView
{
    int32_t nB;
    const int32_t* faceCells;   // [nB]
    const int32_t* cellStart;   // [nCells+1] natural numbering
    const int32_t* cellFaces;   // [nB]
};

struct LaplacianInputs
{
    const double *gamma, *weights, *magSf, *deltaCoeffs;
    const double *bMagSf, *bDeltaCoeffs, *bValue;
    const int32_t* bType;
};

struct ConvDiffInputs
{
    const double* phi;
    double D;
    const double *magSf, *deltaCoeffs;
    const double *bPhi, *bMagSf, *bDeltaCoeffs, *bValue;
    const int32_t* bType;
};

struct GradInputs
{
    const double *phi, *weights, *Sf, *V;
    const double *bSf, *bMagSf, *bPhi, *bSnGrad;
};

void launch_laplacian(const RowsView& R, const int32_t* perm, const int32_t* lower,
                      const int32_t* upper, int32_t F, const BoundaryView& B,
                      const LaplacianInputs& in, double* upperF, double* diagLib,
                      double* ic, double* bc, double* source, int32_t addBoundaryDiag,
                      hipStream_t s);
void launch_convection_diffusion(const RowsView& R, const int32_t* perm, int32_t F,
                                 const BoundaryView& B, const ConvDiffInputs& in,
                                 double* upperF, double* lowerF, double* work3F,
                                 double* diagLib, double* ic, double* bc, double* source,
                                 int32_t addBoundaryDiag, hipStream_t s);
void launch_grad(const RowsView& R, const int32_t* perm, const int32_t* lower,
                 const int32_t* upper, int32_t F, const BoundaryView& B, const GradInputs& in,
                 double* work3F, double* grad, double* bGrad, hipStream_t s);

// mesh geometry (fh_geometry.hip)
void launch_face_geometry(int32_t nFaces, const int32_t* faceStart, const int32_t* facePoints,
                          const double* points, double* Cf, double* Sf, double* magSf,
                          hipStream_t s);
void launch_cell_geometry(const RowsView& R, const int32_t* perm, const BoundaryView& B,
                          const double* Cf, const double* Sf, const double* bCf,
                          const double* bSf, double* C, double* V, hipStream_t s);
void launch_surface_interpolation(int32_t F, const int32_t* lower, const int32_t* upper,
                                  const double* C, const double* Cf, const double* Sf,
                                  const double* magSf, double* w, double* dc, double* ndc,
                                  int32_t nB, const int32_t* faceCells, const double* bCf,
                                  const double* bSf, const double* bMagSf, double* bdc,
                                  double* bndc, hipStream_t s);

} // namespace fh
