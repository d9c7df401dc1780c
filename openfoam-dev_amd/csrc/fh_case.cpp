// Synthetic cases: hex boxes in blockMesh order with their finite-volume
// geometry, and the seeded fields of the measured configurations.  Host
// code only (libfoamhip_case.so); it produces inputs, it is not on the
// solve path.
//
// Ordering (blockCreate.C:349-377, polyMeshFromShapeMesh.C:170-281): cell
// i + nx*(j + ny*k); internal faces per owner in increasing neighbour order
// (x, then y, then z neighbour).  Boundary patches: 0 x-min, 1 x-max,
// 2 y-min, 3 y-max, 4 z-min, 5 z-max, faces in cell order within a patch.
//
// Random numbers: counter-based SplitMix64, u(seed, i) in [0, 1).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace
{

inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27))*0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double u01(uint64_t seed, uint64_t i)
{
    const uint64_t z = splitmix64(splitmix64(seed) ^ (i*0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
    return (double)(z >> 11)*(1.0/9007199254740992.0);
}

inline double normal01(uint64_t seed, uint64_t i)
{
    // Box-Muller on two counter draws
    double u1 = u01(seed, 2*i);
    const double u2 = u01(seed, 2*i + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0*std::log(u1))*std::cos(2.0*M_PI*u2);
}

} // namespace

extern "C" {

int64_t fhc_box_nfaces(int32_t nx, int32_t ny, int32_t nz)
{
    return (int64_t)(nx - 1)*ny*nz + (int64_t)nx*(ny - 1)*nz + (int64_t)nx*ny*(nz - 1);
}

int64_t fhc_box_nbfaces(int32_t nx, int32_t ny, int32_t nz)
{
    return 2*((int64_t)ny*nz + (int64_t)nx*nz + (int64_t)nx*ny);
}

// internal-face addressing; optional per-face direction (0 x, 1 y, 2 z)
void fhc_box_faces(int32_t nx, int32_t ny, int32_t nz, int32_t* lower, int32_t* upper,
                   int8_t* dir)
{
    int64_t f = 0;
    for (int32_t k = 0; k < nz; k++)
        for (int32_t j = 0; j < ny; j++)
            for (int32_t i = 0; i < nx; i++)
            {
                const int32_t c = i + nx*(j + ny*k);
                if (i < nx - 1) { lower[f] = c; upper[f] = c + 1; if (dir) dir[f] = 0; f++; }
                if (j < ny - 1) { lower[f] = c; upper[f] = c + nx; if (dir) dir[f] = 1; f++; }
                if (k < nz - 1) { lower[f] = c; upper[f] = c + nx*ny; if (dir) dir[f] = 2; f++; }
            }
}

// boundary faces: cell, patch id and outward unit normal direction
// (axis 0..2, sign +-1)
void fhc_box_bfaces(int32_t nx, int32_t ny, int32_t nz, int32_t* faceCells, int8_t* patch)
{
    int64_t b = 0;
    for (int p = 0; p < 6; p++)
    {
        const int axis = p/2;
        const bool hi = p & 1;
        for (int32_t k = 0; k < nz; k++)
            for (int32_t j = 0; j < ny; j++)
                for (int32_t i = 0; i < nx; i++)
                {
                    const int32_t idx[3] = {i, j, k};
                    const int32_t nmax[3] = {nx, ny, nz};
                    if (idx[axis] != (hi ? nmax[axis] - 1 : 0)) continue;
                    faceCells[b] = i + nx*(j + ny*k);
                    patch[b] = (int8_t)p;
                    b++;
                }
    }
}

// Geometry of a uniform box [0,Lx]x[0,Ly]x[0,Lz]: the values OpenFOAM's
// surfaceInterpolation (surfaceInterpolation.C:135-330) and fvPatch::delta
// (fvPatch.C:159-164) give for orthogonal uniform hexes: weights 1/2,
// deltaCoeffs = nonOrthDeltaCoeffs = 1/h, boundary deltaCoeffs 2/h.
void fhc_box_geometry(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz,
                      double* weights, double* magSf, double* Sf, double* deltaCoeffs,
                      double* V, double* C, double* Cf,
                      double* bMagSf, double* bSf, double* bDeltaCoeffs, double* bCf)
{
    const double h[3] = {Lx/nx, Ly/ny, Lz/nz};
    const double A[3] = {h[1]*h[2], h[0]*h[2], h[0]*h[1]};
    int64_t f = 0;
    for (int32_t k = 0; k < nz; k++)
        for (int32_t j = 0; j < ny; j++)
            for (int32_t i = 0; i < nx; i++)
            {
                const int32_t c = i + nx*(j + ny*k);
                const double cc[3] = {(i + 0.5)*h[0], (j + 0.5)*h[1], (k + 0.5)*h[2]};
                if (V) V[c] = h[0]*h[1]*h[2];
                if (C) for (int d = 0; d < 3; d++) C[3*c + d] = cc[d];
                const int32_t idx[3] = {i, j, k};
                const int32_t nmax[3] = {nx, ny, nz};
                for (int d = 0; d < 3; d++)
                {
                    if (idx[d] >= nmax[d] - 1) continue;
                    if (weights) weights[f] = 0.5;
                    if (magSf) magSf[f] = A[d];
                    if (Sf) { Sf[3*f] = Sf[3*f + 1] = Sf[3*f + 2] = 0.0; Sf[3*f + d] = A[d]; }
                    if (deltaCoeffs) deltaCoeffs[f] = 1.0/h[d];
                    if (Cf)
                    {
                        for (int e = 0; e < 3; e++) Cf[3*f + e] = cc[e];
                        Cf[3*f + d] = (idx[d] + 1)*h[d];
                    }
                    f++;
                }
            }
    int64_t b = 0;
    for (int p = 0; p < 6; p++)
    {
        const int axis = p/2;
        const bool hi = p & 1;
        for (int32_t k = 0; k < nz; k++)
            for (int32_t j = 0; j < ny; j++)
                for (int32_t i = 0; i < nx; i++)
                {
                    const int32_t idx[3] = {i, j, k};
                    const int32_t nmax[3] = {nx, ny, nz};
                    if (idx[axis] != (hi ? nmax[axis] - 1 : 0)) continue;
                    if (bMagSf) bMagSf[b] = A[axis];
                    if (bSf)
                    {
                        bSf[3*b] = bSf[3*b + 1] = bSf[3*b + 2] = 0.0;
                        bSf[3*b + axis] = hi ? A[axis] : -A[axis];
                    }
                    if (bDeltaCoeffs) bDeltaCoeffs[b] = 2.0/h[axis];
                    if (bCf)
                    {
                        const double cc[3] = {(i + 0.5)*h[0], (j + 0.5)*h[1], (k + 0.5)*h[2]};
                        for (int e = 0; e < 3; e++) bCf[3*b + e] = cc[e];
                        bCf[3*b + axis] = hi ? (idx[axis] + 1)*h[axis] : 0.0;
                    }
                    b++;
                }
    }
}

// n values uniform in [lo, hi) from counter draws (seed, offset + i)
void fhc_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = lo + (hi - lo)*u01(seed, (uint64_t)i);
}

// zero-mean uniform right-hand side: U[-1,1] minus its mean
void fhc_rhs(uint64_t seed, int64_t n, double* out)
{
    fhc_uniform(seed, n, -1.0, 1.0, out);
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += out[i];
    const double m = s/n;
    for (int64_t i = 0; i < n; i++) out[i] -= m;
}

// log-normal diffusivity gamma = exp(sigma*G): G is seeded white noise
// convolved with a Gaussian kernel exp(-d^2/(2 l^2)) (|d| <= 3l, zero
// padding), then normalised to zero mean and unit variance.
void fhc_lognormal(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double corrLen,
                   double sigma, double* gamma)
{
    const int64_t N = (int64_t)nx*ny*nz;
    std::vector<double> a(N), b(N);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) a[i] = normal01(seed, (uint64_t)i);
    const int R = corrLen > 0 ? (int)std::ceil(3.0*corrLen) : 0;
    std::vector<double> ker(2*R + 1);
    for (int d = -R; d <= R; d++)
        ker[d + R] = corrLen > 0 ? std::exp(-0.5*d*d/(corrLen*corrLen)) : 1.0;
    const int32_t n3[3] = {nx, ny, nz};
    const int64_t stride[3] = {1, nx, (int64_t)nx*ny};
    for (int axis = 0; axis < 3 && R > 0; axis++)
    {
        const int32_t L = n3[axis];
        const int64_t s = stride[axis];
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < N; c++)
        {
            const int32_t pos = (int32_t)((c/s) % L);
            double acc = 0.0;
            for (int d = -R; d <= R; d++)
            {
                const int32_t q = pos + d;
                if (q < 0 || q >= L) continue;
                acc += ker[d + R]*a[c + (int64_t)d*s];
            }
            b[c] = acc;
        }
        a.swap(b);
    }
    double mean = 0.0;
    for (int64_t i = 0; i < N; i++) mean += a[i];
    mean /= N;
    double var = 0.0;
    for (int64_t i = 0; i < N; i++) var += (a[i] - mean)*(a[i] - mean);
    const double sd = std::sqrt(var/N);
    const double inv = sd > 0 ? 1.0/sd : 1.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) gamma[i] = std::exp(sigma*(a[i] - mean)*inv);
}

// Divergence-free velocity U(x) = sum_m (k_m x a_m) sin(2 pi k_m.x + phi_m),
// k_m integer in [1,4]^3, a_m in [-1,1]^3, phi_m in [0, 2 pi), m = 0..7,
// scaled to unit RMS |U| over the cell centres.  Returns the face fluxes
// phi = U(Cf).Sf on internal and boundary faces.
void fhc_fourier_flux(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, const double* C,
                      const double* Cf, const double* Sf, const double* bCf,
                      const double* bSf, double* phi, double* bPhi)
{
    double K[8][3], Aa[8][3], P[8];
    uint64_t q = 0;
    for (int m = 0; m < 8; m++)
    {
        for (int d = 0; d < 3; d++) K[m][d] = 1.0 + std::floor(4.0*u01(seed, q++));
        for (int d = 0; d < 3; d++) Aa[m][d] = -1.0 + 2.0*u01(seed, q++);
        P[m] = 2.0*M_PI*u01(seed, q++);
    }
    double kxa[8][3];
    for (int m = 0; m < 8; m++)
    {
        kxa[m][0] = K[m][1]*Aa[m][2] - K[m][2]*Aa[m][1];
        kxa[m][1] = K[m][2]*Aa[m][0] - K[m][0]*Aa[m][2];
        kxa[m][2] = K[m][0]*Aa[m][1] - K[m][1]*Aa[m][0];
    }
    auto U = [&](const double* x, double* u)
    {
        u[0] = u[1] = u[2] = 0.0;
        for (int m = 0; m < 8; m++)
        {
            const double s = std::sin(2.0*M_PI*(K[m][0]*x[0] + K[m][1]*x[1] + K[m][2]*x[2]) + P[m]);
            for (int d = 0; d < 3; d++) u[d] += kxa[m][d]*s;
        }
    };
    const int64_t N = (int64_t)nx*ny*nz;
    double ss = 0.0;
#pragma omp parallel for reduction(+:ss) schedule(static)
    for (int64_t c = 0; c < N; c++)
    {
        double u[3];
        U(C + 3*c, u);
        ss += u[0]*u[0] + u[1]*u[1] + u[2]*u[2];
    }
    const double scale = ss > 0 ? 1.0/std::sqrt(ss/N) : 1.0;
    const int64_t F = (int64_t)(nx - 1)*ny*nz + (int64_t)nx*(ny - 1)*nz + (int64_t)nx*ny*(nz - 1);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; f++)
    {
        double u[3];
        U(Cf + 3*f, u);
        phi[f] = scale*(u[0]*Sf[3*f] + u[1]*Sf[3*f + 1] + u[2]*Sf[3*f + 2]);
    }
    const int64_t B = 2*((int64_t)ny*nz + (int64_t)nx*nz + (int64_t)nx*ny);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < B; b++)
    {
        double u[3];
        U(bCf + 3*b, u);
        bPhi[b] = scale*(u[0]*bSf[3*b] + u[1]*bSf[3*b + 1] + u[2]*bSf[3*b + 2]);
    }
}

} // extern "C"

extern "C" {

// Points of a box [0,Lx]x[0,Ly]x[0,Lz] with nx x ny x nz hex cells, graded
// along z with expansion ratio zRatio (last cell over first, blockMesh's
// simpleGrading (1 1 zRatio): sizes in geometric progression with ratio
// zRatio^(1/(nz-1))), then jittered: every coordinate of a point that is not
// on a boundary plane normal to it moves by jitter * (local cell size along
// that axis, the smaller of the two cells it separates) * U[-1, 1) (seeded).
// Point P(i, j, k) = i + (nx+1)*(j + (ny+1)*k); points[3p + axis].
void fhc_graded_box_points(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz,
                           double zRatio, double jitter, uint64_t seed, double* points)
{
    std::vector<double> zs(nz + 1);
    if (nz > 1 && zRatio != 1.0)
    {
        const double r = std::pow(zRatio, 1.0/(nz - 1));
        for (int32_t k = 0; k <= nz; k++) zs[k] = Lz*(1.0 - std::pow(r, k))/(1.0 - std::pow(r, nz));
    }
    else
        for (int32_t k = 0; k <= nz; k++) zs[k] = Lz*k/nz;
    zs[nz] = Lz;
    const double hx = Lx/nx, hy = Ly/ny;
    for (int32_t k = 0; k <= nz; k++)
        for (int32_t j = 0; j <= ny; j++)
            for (int32_t i = 0; i <= nx; i++)
            {
                const int64_t p = i + (int64_t)(nx + 1)*(j + (int64_t)(ny + 1)*k);
                double x = hx*i, y = hy*j, z = zs[k];
                if (i > 0 && i < nx) x += jitter*hx*(2.0*u01(seed, 3*p) - 1.0);
                if (j > 0 && j < ny) y += jitter*hy*(2.0*u01(seed, 3*p + 1) - 1.0);
                if (k > 0 && k < nz)
                {
                    const double hz = std::min(zs[k] - zs[k - 1], zs[k + 1] - zs[k]);
                    z += jitter*hz*(2.0*u01(seed, 3*p + 2) - 1.0);
                }
                points[3*p] = x;
                points[3*p + 1] = y;
                points[3*p + 2] = z;
            }
}

// The faces of that box as quadrilaterals, in the order of fhc_box_faces
// (internal) followed by fhc_box_bfaces (boundary): facePoints[4f .. 4f+3],
// ordered so that the face normal points from owner to neighbour (internal)
// or out of the box (boundary), as polyMesh orients them.
void fhc_box_face_points(int32_t nx, int32_t ny, int32_t nz, int32_t* facePoints)
{
    auto P = [&](int32_t i, int32_t j, int32_t k) {
        return (int32_t)(i + (int64_t)(nx + 1)*(j + (int64_t)(ny + 1)*k));
    };
    // a face on the plane normal to axis at index (i, j, k) of its low
    // corner, normal +axis (reversed for -axis)
    auto quad = [&](int axis, int32_t i, int32_t j, int32_t k, bool neg, int32_t* out) {
        int32_t q[4];
        if (axis == 0)
        {
            q[0] = P(i, j, k); q[1] = P(i, j + 1, k); q[2] = P(i, j + 1, k + 1); q[3] = P(i, j, k + 1);
        }
        else if (axis == 1)
        {
            q[0] = P(i, j, k); q[1] = P(i, j, k + 1); q[2] = P(i + 1, j, k + 1); q[3] = P(i + 1, j, k);
        }
        else
        {
            q[0] = P(i, j, k); q[1] = P(i + 1, j, k); q[2] = P(i + 1, j + 1, k); q[3] = P(i, j + 1, k);
        }
        if (neg) { out[0] = q[0]; out[1] = q[3]; out[2] = q[2]; out[3] = q[1]; }
        else for (int m = 0; m < 4; m++) out[m] = q[m];
    };
    int64_t f = 0;
    for (int32_t k = 0; k < nz; k++)
        for (int32_t j = 0; j < ny; j++)
            for (int32_t i = 0; i < nx; i++)
            {
                if (i < nx - 1) quad(0, i + 1, j, k, false, facePoints + 4*f++);
                if (j < ny - 1) quad(1, i, j + 1, k, false, facePoints + 4*f++);
                if (k < nz - 1) quad(2, i, j, k + 1, false, facePoints + 4*f++);
            }
    for (int p = 0; p < 6; p++)
    {
        const int axis = p/2;
        const bool hi = p & 1;
        for (int32_t k = 0; k < nz; k++)
            for (int32_t j = 0; j < ny; j++)
                for (int32_t i = 0; i < nx; i++)
                {
                    const int32_t idx[3] = {i, j, k};
                    const int32_t nmax[3] = {nx, ny, nz};
                    if (idx[axis] != (hi ? nmax[axis] - 1 : 0)) continue;
                    int32_t c[3] = {i, j, k};
                    if (hi) c[axis] += 1;
                    quad(axis, c[0], c[1], c[2], !hi, facePoints + 4*f++);
                }
    }
}

} // extern "C"
