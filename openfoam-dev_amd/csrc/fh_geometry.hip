// Mesh geometry on the device, as the reference computes it on the host:
//   face centres and area vectors   face::areaAndCentre (faceTemplates.C:79-149),
//                                   magSf = max(mag(Sf), rootVSmall)
//                                   (primitiveMeshFaceCentresAndAreas.C)
//   cell centres and volumes        primitiveMesh::makeCellCentresAndVols
//                                   (primitiveMeshCellCentresAndVols.C:72-154)
//   interpolation weights and       surfaceInterpolation::makeWeights,
//   delta coefficients              makeDeltaCoeffs, makeNonOrthDeltaCoeffs
//                                   (surfaceInterpolation.C:135-330), fvPatch
//                                   delta / nf (fvPatch.C:138-164)
// One thread per face for the face loops; the cell loop is a gather: each
// cell takes its owner faces (internal in face order, then boundary in
// patch order) and then its neighbour faces, the order of the reference's
// scattered loops, so every value is the reference's to the bit (the
// library is built without FMA contraction).
#include "fh_assembly.h"

#include <cfloat>

namespace fh
{

#define GE_B 256

struct V3 { double x, y, z; };

__device__ __forceinline__ V3 ld3(const double* a, int64_t i) { return {a[3*i], a[3*i + 1], a[3*i + 2]}; }
__device__ __forceinline__ void st3(double* a, int64_t i, V3 v) { a[3*i] = v.x; a[3*i + 1] = v.y; a[3*i + 2] = v.z; }
__device__ __forceinline__ V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul3(double s, V3 a) { return {s*a.x, s*a.y, s*a.z}; }
__device__ __forceinline__ V3 div3(V3 a, double s) { return {a.x/s, a.y/s, a.z/s}; }
// Vector::operator& and operator^ (VectorI.H:145-160); mag = sqrt(x^2 + y^2 + z^2)
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x*b.x + a.y*b.y + a.z*b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b)
{
    return {a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x};
}
__device__ __forceinline__ double mag3(V3 a) { return sqrt(a.x*a.x + a.y*a.y + a.z*a.z); }

// face::areaAndCentre: a triangle directly; otherwise the fan from the
// point average, centres weighted by the triangle areas projected on the
// face normal
__global__ void k_face_geometry(int32_t nFaces, const int32_t* __restrict__ faceStart,
                                const int32_t* __restrict__ facePoints,
                                const double* __restrict__ points, double* __restrict__ Cf,
                                double* __restrict__ Sf, double* __restrict__ magSf)
{
    const int32_t f = blockIdx.x*GE_B + threadIdx.x;
    if (f >= nFaces) return;
    const int32_t s = faceStart[f], n = faceStart[f + 1] - s;
    const int32_t* pts = facePoints + s;
    V3 area, centre;
    if (n == 3)
    {
        const V3 p0 = ld3(points, pts[0]), p1 = ld3(points, pts[1]), p2 = ld3(points, pts[2]);
        area = mul3(1.0/2.0, cross3(sub3(p1, p0), sub3(p2, p0)));
        centre = mul3(1.0/3.0, add3(add3(p0, p1), p2));
    }
    else
    {
        V3 pAvg = {0.0, 0.0, 0.0};
        for (int32_t i = 0; i < n; i++) pAvg = add3(pAvg, ld3(points, pts[i]));
        pAvg = div3(pAvg, (double)n);
        V3 sumA = {0.0, 0.0, 0.0};
        for (int32_t i = 0; i < n; i++)
        {
            const V3 p = ld3(points, pts[i]), pn = ld3(points, pts[i + 1 < n ? i + 1 : 0]);
            sumA = add3(sumA, cross3(sub3(pn, p), sub3(pAvg, p)));
        }
        const double mA = mag3(sumA);
        const V3 hat = mA > 0 ? div3(sumA, mA) : V3{0.0, 0.0, 0.0};
        double sumAn = 0.0;
        V3 sumAnc = {0.0, 0.0, 0.0};
        for (int32_t i = 0; i < n; i++)
        {
            const V3 p = ld3(points, pts[i]), pn = ld3(points, pts[i + 1 < n ? i + 1 : 0]);
            const V3 a = cross3(sub3(pn, p), sub3(pAvg, p));
            const V3 c = add3(add3(p, pn), pAvg);
            const double an = dot3(a, hat);
            sumAn += an;
            sumAnc = add3(sumAnc, mul3(an, c));
        }
        area = mul3(1.0/2.0, sumA);
        centre = sumAn > DBL_MIN ? div3(mul3(1.0/3.0, sumAnc), sumAn) : pAvg;
    }
    st3(Cf, f, centre);
    st3(Sf, f, area);
    const double m = mag3(area);
    const double rootVSmall = 1.4916681462400413e-154;     // sqrt(DBL_MIN)
    magSf[f] = m > rootVSmall ? m : rootVSmall;
}

// the face centre / area of the row's face entry k and its kind
struct CellFace { V3 c, a; };

__device__ __forceinline__ int32_t ge_base(const int32_t* sliceStart, int32_t r)
{
    return sliceStart[r >> 6] + (r & 63);
}

// makeCellCentresAndVols as a gather per library row (cell perm[r])
__global__ void k_cell_geometry(RowsView R, const int32_t* __restrict__ perm, BoundaryView B,
                                const double* __restrict__ Cf, const double* __restrict__ Sf,
                                const double* __restrict__ bCf, const double* __restrict__ bSf,
                                double* __restrict__ C, double* __restrict__ V)
{
    const int32_t r = blockIdx.x*GE_B + threadIdx.x;
    if (r >= R.n) return;
    const int32_t cell = perm[r];
    const int32_t info = R.rowInfo[r];
    const int32_t nLower = info & 0xffff, nEnt = info >> 16;
    const int32_t base = ge_base(R.sliceStart, r);
    const int32_t b0 = B.nB > 0 ? B.cellStart[cell] : 0, b1 = B.nB > 0 ? B.cellStart[cell + 1] : 0;
    // owner faces: the row's upper entries (internal, face order), then its
    // boundary faces; neighbour faces: the row's lower entries
    V3 cEst = {0.0, 0.0, 0.0};
    int32_t nF = 0;
    for (int32_t k = nLower; k < nEnt; k++, nF++)
        cEst = add3(cEst, ld3(Cf, R.face[base + (k << 6)] >> 1));
    for (int32_t q = b0; q < b1; q++, nF++) cEst = add3(cEst, ld3(bCf, B.cellFaces[q]));
    for (int32_t k = 0; k < nLower; k++, nF++)
        cEst = add3(cEst, ld3(Cf, R.face[base + (k << 6)] >> 1));
    cEst = div3(cEst, (double)nF);
    V3 cc = {0.0, 0.0, 0.0};
    double vol = 0.0;
    for (int32_t k = nLower; k < nEnt; k++)
    {
        const int32_t f = R.face[base + (k << 6)] >> 1;
        const V3 fc = ld3(Cf, f), fa = ld3(Sf, f);
        const double pyr3Vol = dot3(fa, sub3(fc, cEst));
        const V3 pc = add3(mul3(3.0/4.0, fc), mul3(1.0/4.0, cEst));
        cc = add3(cc, mul3(pyr3Vol, pc));
        vol += pyr3Vol;
    }
    for (int32_t q = b0; q < b1; q++)
    {
        const int32_t b = B.cellFaces[q];
        const V3 fc = ld3(bCf, b), fa = ld3(bSf, b);
        const double pyr3Vol = dot3(fa, sub3(fc, cEst));
        const V3 pc = add3(mul3(3.0/4.0, fc), mul3(1.0/4.0, cEst));
        cc = add3(cc, mul3(pyr3Vol, pc));
        vol += pyr3Vol;
    }
    for (int32_t k = 0; k < nLower; k++)
    {
        const int32_t f = R.face[base + (k << 6)] >> 1;
        const V3 fc = ld3(Cf, f), fa = ld3(Sf, f);
        const double pyr3Vol = dot3(fa, sub3(cEst, fc));
        const V3 pc = add3(mul3(3.0/4.0, fc), mul3(1.0/4.0, cEst));
        cc = add3(cc, mul3(pyr3Vol, pc));
        vol += pyr3Vol;
    }
    st3(C, cell, fabs(vol) > DBL_MIN ? div3(cc, vol) : cEst);
    V[cell] = vol*(1.0/3.0);
}

// makeWeights, makeDeltaCoeffs, makeNonOrthDeltaCoeffs of the internal faces
__global__ void k_surface_interpolation(int32_t F, const int32_t* __restrict__ lower,
                                        const int32_t* __restrict__ upper,
                                        const double* __restrict__ C,
                                        const double* __restrict__ Cf,
                                        const double* __restrict__ Sf,
                                        const double* __restrict__ magSf,
                                        double* __restrict__ w, double* __restrict__ dc,
                                        double* __restrict__ ndc)
{
    const int32_t f = blockIdx.x*GE_B + threadIdx.x;
    if (f >= F) return;
    const V3 sf = ld3(Sf, f), cf = ld3(Cf, f);
    const V3 cO = ld3(C, lower[f]), cN = ld3(C, upper[f]);
    const double SfdOwn = fabs(dot3(sf, sub3(cf, cO)));
    const double SfdNei = fabs(dot3(sf, sub3(cN, cf)));
    const double SfdOwnNei = SfdOwn + SfdNei;
    const double vGreat = DBL_MAX/10;
    if (SfdNei/vGreat < SfdOwnNei) w[f] = SfdNei/SfdOwnNei;
    else
    {
        const double dOwn = mag3(sub3(cf, cO)), dNei = mag3(sub3(cN, cf));
        w[f] = dNei/(dOwn + dNei);
    }
    const V3 delta = sub3(cN, cO);
    dc[f] = 1.0/mag3(delta);
    const V3 unitArea = div3(sf, magSf[f]);
    const double a = dot3(unitArea, delta), b = 0.05*mag3(delta);
    ndc[f] = 1.0/(a > b ? a : b);
}

// non-coupled patch faces: delta = nf (nf & (Cf - Cn))
__global__ void k_patch_delta_coeffs(int32_t nB, const int32_t* __restrict__ faceCells,
                                     const double* __restrict__ C,
                                     const double* __restrict__ bCf,
                                     const double* __restrict__ bSf,
                                     const double* __restrict__ bMagSf,
                                     double* __restrict__ bdc, double* __restrict__ bndc)
{
    const int32_t b = blockIdx.x*GE_B + threadIdx.x;
    if (b >= nB) return;
    const V3 nf = div3(ld3(bSf, b), bMagSf[b]);
    const V3 d = sub3(ld3(bCf, b), ld3(C, faceCells[b]));
    const V3 delta = mul3(dot3(nf, d), nf);
    bdc[b] = 1.0/mag3(delta);
    if (bndc)
    {
        const double a = dot3(nf, delta), c = 0.05*mag3(delta);
        bndc[b] = 1.0/(a > c ? a : c);
    }
}

void launch_face_geometry(int32_t nFaces, const int32_t* faceStart, const int32_t* facePoints,
                          const double* points, double* Cf, double* Sf, double* magSf,
                          hipStream_t s)
{
    if (nFaces <= 0) return;
    hipLaunchKernelGGL(k_face_geometry, dim3((nFaces + GE_B - 1)/GE_B), dim3(GE_B), 0, s, nFaces,
                       faceStart, facePoints, points, Cf, Sf, magSf);
}

void launch_cell_geometry(const RowsView& R, const int32_t* perm, const BoundaryView& B,
                          const double* Cf, const double* Sf, const double* bCf,
                          const double* bSf, double* C, double* V, hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_geometry, dim3((R.n + GE_B - 1)/GE_B), dim3(GE_B), 0, s, R, perm,
                       B, Cf, Sf, bCf, bSf, C, V);
}

void launch_surface_interpolation(int32_t F, const int32_t* lower, const int32_t* upper,
                                  const double* C, const double* Cf, const double* Sf,
                                  const double* magSf, double* w, double* dc, double* ndc,
                                  int32_t nB, const int32_t* faceCells, const double* bCf,
                                  const double* bSf, const double* bMagSf, double* bdc,
                                  double* bndc, hipStream_t s)
{
    if (F > 0)
        hipLaunchKernelGGL(k_surface_interpolation, dim3((F + GE_B - 1)/GE_B), dim3(GE_B), 0, s,
                           F, lower, upper, C, Cf, Sf, magSf, w, dc, ndc);
    if (nB > 0)
        hipLaunchKernelGGL(k_patch_delta_coeffs, dim3((nB + GE_B - 1)/GE_B), dim3(GE_B), 0, s,
                           nB, faceCells, C, bCf, bSf, bMagSf, bdc, bndc);
}

} // namespace fh
