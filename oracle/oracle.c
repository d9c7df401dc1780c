/*
 * oracle.c -- CPU restatement of OpenFOAM-dev's lduMatrix solver path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the
 * reference routine whose loop order it follows.  Build with
 * -O3 -ffp-contract=off (no FMA contraction, no fast-math), matching
 * wmake's linux64Gcc Opt rules (wmake/rules/linux64Gcc/c++Opt).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_GREAT_ 1e20          /* SolverPerformance.H:285 */
#define OR_SMALL_ 1e-20         /* SolverPerformance.H:288 */
#define OR_VSMALL_ DBL_MIN      /* SolverPerformance.H:291, doubleScalar.H:57 */
#define OR_SMALL DBL_EPSILON    /* doubleScalar.H:60 */
#define OR_GREAT (1.0/DBL_EPSILON)

static double* dalloc(int32_t n) { return (double*)calloc(n > 0 ? n : 1, sizeof(double)); }
static int32_t* ialloc(int32_t n) { return (int32_t*)calloc(n > 0 ? n : 1, sizeof(int32_t)); }

/* ------------------------------------------------------------------ */
/* addressing: lduAddressing.C:32-170                                  */
/* ------------------------------------------------------------------ */

int or_addr_init(or_addr* a, int32_t nCells, int32_t nFaces,
                 const int32_t* lower, const int32_t* upper)
{
    a->nCells = nCells;
    a->nFaces = nFaces;
    a->lower = lower;
    a->upper = upper;
    a->losort = ialloc(nFaces);
    a->ownerStart = ialloc(nCells + 1);
    a->losortStart = ialloc(nCells + 1);

    /* calcLosort (lduAddressing.C:32-90): faces grouped by neighbour,
       ascending face index within a group */
    int32_t* nNbr = ialloc(nCells);
    for (int32_t f = 0; f < nFaces; f++) nNbr[upper[f]]++;
    int32_t* off = ialloc(nCells + 1);
    for (int32_t c = 0; c < nCells; c++) off[c + 1] = off[c] + nNbr[c];
    memset(nNbr, 0, sizeof(int32_t) * (nCells > 0 ? nCells : 1));
    for (int32_t f = 0; f < nFaces; f++)
    {
        int32_t u = upper[f];
        a->losort[off[u] + nNbr[u]] = f;
        nNbr[u]++;
    }

    /* calcOwnerStart (lduAddressing.C:93-127) */
    for (int32_t c = 0; c <= nCells; c++) a->ownerStart[c] = nFaces;
    a->ownerStart[0] = 0;
    {
        int32_t nOwnStart = 0, i = 1;
        for (int32_t f = 0; f < nFaces; f++)
        {
            int32_t curOwn = lower[f];
            if (curOwn > nOwnStart)
            {
                while (i <= curOwn) a->ownerStart[i++] = f;
                nOwnStart = curOwn;
            }
        }
    }

    /* calcLosortStart (lduAddressing.C:130-170) */
    for (int32_t c = 0; c <= nCells; c++) a->losortStart[c] = 0;
    {
        int32_t nLsrtStart = 0, i = 0;
        for (int32_t f = 0; f < nFaces; f++)
        {
            int32_t curNbr = upper[a->losort[f]];
            if (curNbr > nLsrtStart)
            {
                while (i <= curNbr) a->losortStart[i++] = f;
                nLsrtStart = curNbr;
            }
        }
        a->losortStart[nCells] = nFaces;
    }

    free(nNbr);
    free(off);
    return 0;
}

void or_addr_free(or_addr* a)
{
    free(a->losort);
    free(a->ownerStart);
    free(a->losortStart);
    a->losort = a->ownerStart = a->losortStart = NULL;
}

/* ------------------------------------------------------------------ */
/* reductions: FieldReductionFunctions.C:113-128, 191-216, 257-290,    */
/* 334-357 (serial left-to-right loops; reduce() is a no-op serially)  */
/* ------------------------------------------------------------------ */

/* Summation order of the reductions: 0 (the reference's serial loop), or 1
   (sums of consecutive blocks of 256 values, then the block sums in order:
   the same algorithm with another order of additions, as a measure of how
   far the reference's results move with the order of the global sums --
   test infrastructure for setting parity bars, never the reference). */
static int g_sum_order = 0;

void or_set_sum_order(int order) { g_sum_order = order; }

#define OR_SUM_LOOP(EXPR)                                               \
    double s = 0.0;                                                     \
    if (g_sum_order == 0)                                               \
    {                                                                   \
        for (int32_t i = 0; i < n; i++) s += (EXPR);                    \
        return s;                                                       \
    }                                                                   \
    for (int32_t i0 = 0; i0 < n; i0 += 256)                             \
    {                                                                   \
        double b = 0.0;                                                 \
        const int32_t i1 = i0 + 256 < n ? i0 + 256 : n;                 \
        for (int32_t i = i0; i < i1; i++) b += (EXPR);                  \
        s += b;                                                         \
    }                                                                   \
    return s;

double or_sum(int32_t n, const double* f) { OR_SUM_LOOP(f[i]) }

double or_sumMag(int32_t n, const double* f) { OR_SUM_LOOP(fabs(f[i])) }

double or_sumSqr(int32_t n, const double* f) { OR_SUM_LOOP(f[i]*f[i]) }

double or_sumProd(int32_t n, const double* a, const double* b_) { OR_SUM_LOOP(a[i]*b_[i]) }

double or_average(int32_t n, const double* f)
{
    if (n <= 0) return 0.0;
    return or_sum(n, f)/n;
}

/* ------------------------------------------------------------------ */
/* matrix kernels                                                      */
/* ------------------------------------------------------------------ */

/* lduMatrix::updateMatrixInterfaces, blocking communication
   (lduMatrixUpdateMatrixInterfaces.C:97-116): every interface in order
   calls its updateInterfaceMatrix, for a cyclic patch
   (cyclicFvPatchField.C:150-174) result[faceCells] -= coeffs*psi[nbr];
   negate: the coefficients enter negated (mBouCoeffs of residual and of
   the Gauss-Seidel smoothers, lduMatrixATmul.C:230-244,
   GaussSeidelSmoother.C:118-125) */
static void update_interfaces(const or_matrix* m, const double* coeffs, int negate,
                              const double* psi, double* result)
{
    const or_interfaces* I = m->interfaces;
    if (!I) return;
    for (int32_t i = 0; i < I->nInterfaces; i++)
        for (int32_t e = I->start[i]; e < I->start[i + 1]; e++)
        {
            const double c = negate ? -coeffs[e] : coeffs[e];
            result[I->faceCells[e]] -= c*psi[I->nbrCells[e]];
        }
}

/* lduMatrixATmul.C:34-92 */
void or_amul(const or_matrix* m, const double* psi, double* Apsi)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++) Apsi[c] = m->diag[c]*psi[c];
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        Apsi[a->upper[f]] += m->lower[f]*psi[a->lower[f]];
        Apsi[a->lower[f]] += m->upper[f]*psi[a->upper[f]];
    }
    if (m->interfaces) update_interfaces(m, m->interfaces->bouCoeffs, 0, psi, Apsi);
}

/* lduMatrixTemplates.C:33-64: H starts at zero; each face takes the
   lower coefficient times the owner's psi from the neighbour's H and the
   upper coefficient times the neighbour's psi from the owner's H, in face
   order, every component alike */
void or_H(const or_matrix* m, int32_t nCmpt, const double* psi, double* Hpsi)
{
    const or_addr* a = m->addr;
    for (int64_t i = 0; i < (int64_t)a->nCells*nCmpt; i++) Hpsi[i] = 0.0;
    if (!m->upper) return;
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        const int64_t l = (int64_t)a->lower[f]*nCmpt, u = (int64_t)a->upper[f]*nCmpt;
        for (int32_t c = 0; c < nCmpt; c++)
        {
            Hpsi[u + c] -= m->lower[f]*psi[l + c];
            Hpsi[l + c] -= m->upper[f]*psi[u + c];
        }
    }
}

/* lduMatrixATmul.C:298-324 */
void or_H1(const or_matrix* m, double* H1)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++) H1[c] = 0.0;
    if (!m->upper) return;
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        H1[a->upper[f]] -= m->lower[f];
        H1[a->lower[f]] -= m->upper[f];
    }
}

/* lduMatrixTemplates.C:77-111 */
int or_faceH(const or_matrix* m, int32_t nCmpt, const double* psi, double* faceHpsi)
{
    const or_addr* a = m->addr;
    if (!m->upper) return -1;
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        const int64_t l = (int64_t)a->lower[f]*nCmpt, u = (int64_t)a->upper[f]*nCmpt;
        for (int32_t c = 0; c < nCmpt; c++)
            faceHpsi[(int64_t)f*nCmpt + c] = m->upper[f]*psi[u + c] - m->lower[f]*psi[l + c];
    }
    return 0;
}

/* lduMatrixOperations.C:77-97 */
void or_sumMagOffDiag(const or_matrix* m, double* sumOff)
{
    const or_addr* a = m->addr;
    if (!m->upper) return;
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        sumOff[a->upper[f]] += fabs(m->lower[f]);
        sumOff[a->lower[f]] += fabs(m->upper[f]);
    }
}

/* fvMatrix.C:953-1028 (non-coupled patches: NeighbourContrib is the
   boundary coefficients themselves) */
int or_flux(const or_matrix* m, int32_t nCmpt, const double* psi, int32_t nB,
            const int32_t* bFaceCells, const double* ic, const double* bc, const double* corr,
            const double* bCorr, double* flux, double* bFlux)
{
    if (or_faceH(m, nCmpt, psi, flux) != 0) return -1;
    if (corr)
        for (int64_t i = 0; i < (int64_t)m->addr->nFaces*nCmpt; i++) flux[i] += corr[i];
    for (int32_t b = 0; b < nB; b++)
        for (int32_t k = 0; k < nCmpt; k++)
        {
            const int64_t i = (int64_t)b*nCmpt + k;
            const double internalContrib = ic[i]*psi[(int64_t)bFaceCells[b]*nCmpt + k];
            bFlux[i] = internalContrib - bc[i];
            if (bCorr) bFlux[i] += bCorr[i];
        }
    return 0;
}

/* fvMatrix.C:594-740 (non-coupled patches; Foam::max and Foam::min select
   s1 > s2 ? s1 : s2 and s1 < s2 ? s1 : s2; cmptMax/cmptMin fold from the
   first component) */
void or_relax(const or_matrix* m, int32_t nB, const int32_t* bFaceCells, const double* ic,
              int32_t nCmpt, double alpha, const double* psi, double* diag, double* source)
{
    const int32_t N = m->addr->nCells;
    if (alpha <= 0) return;
    double* D0 = (double*)malloc(sizeof(double)*(N > 0 ? N : 1));
    double* sumOff = (double*)calloc(N > 0 ? N : 1, sizeof(double));
    for (int32_t c = 0; c < N; c++) D0[c] = diag[c];
    or_sumMagOffDiag(m, sumOff);
    for (int32_t b = 0; b < nB; b++)
    {
        const double* v = ic + (int64_t)b*nCmpt;
        double mx = fabs(v[0]);
        for (int32_t k = 1; k < nCmpt; k++) { const double a = fabs(v[k]); mx = mx > a ? mx : a; }
        diag[bFaceCells[b]] += mx;
    }
    for (int32_t c = 0; c < N; c++)
    {
        const double a = fabs(diag[c]);
        diag[c] = a > sumOff[c] ? a : sumOff[c];
    }
    for (int32_t c = 0; c < N; c++) diag[c] /= alpha;
    for (int32_t b = 0; b < nB; b++)
    {
        const double* v = ic + (int64_t)b*nCmpt;
        double mn = v[0];
        for (int32_t k = 1; k < nCmpt; k++) mn = mn < v[k] ? mn : v[k];
        diag[bFaceCells[b]] -= mn;
    }
    for (int32_t c = 0; c < N; c++)
    {
        const double dD = diag[c] - D0[c];
        for (int32_t k = 0; k < nCmpt; k++)
            source[(int64_t)c*nCmpt + k] += dD*psi[(int64_t)c*nCmpt + k];
    }
    free(D0);
    free(sumOff);
}

/* lduMatrixATmul.C:95-151 */
void or_tmul(const or_matrix* m, const double* psi, double* Tpsi)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++) Tpsi[c] = m->diag[c]*psi[c];
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        Tpsi[a->upper[f]] += m->upper[f]*psi[a->lower[f]];
        Tpsi[a->lower[f]] += m->lower[f]*psi[a->upper[f]];
    }
    /* lduMatrixATmul.C:132-141: the interfaceIntCoeffs */
    if (m->interfaces) update_interfaces(m, m->interfaces->intCoeffs, 0, psi, Tpsi);
}

/* lduMatrixATmul.C:154-200 */
void or_sumA(const or_matrix* m, double* sumA)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++) sumA[c] = m->diag[c];
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        sumA[a->upper[f]] += m->lower[f];
        sumA[a->lower[f]] += m->upper[f];
    }
    /* lduMatrixATmul.C:185-199: the interface boundary coefficients
       subtracted from the sum, interface by interface in face order */
    const or_interfaces* I = m->interfaces;
    if (I)
        for (int32_t i = 0; i < I->nInterfaces; i++)
            for (int32_t e = I->start[i]; e < I->start[i + 1]; e++)
                sumA[I->faceCells[e]] -= I->bouCoeffs[e];
}

/* lduMatrixATmul.C:203-280 */
void or_residual(const or_matrix* m, const double* psi, const double* source,
                 double* rA)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++)
        rA[c] = source[c] - m->diag[c]*psi[c];
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        rA[a->upper[f]] -= m->lower[f]*psi[a->lower[f]];
        rA[a->lower[f]] -= m->upper[f]*psi[a->upper[f]];
    }
    /* the interfaces with the negated boundary coefficients
       (lduMatrixATmul.C:230-244, 266-273) */
    if (m->interfaces) update_interfaces(m, m->interfaces->bouCoeffs, 1, psi, rA);
}

/* lduMatrixSolver.C:174-197 */
double or_normFactor(const or_matrix* m, const double* psi,
                     const double* source, const double* Apsi, double* tmp)
{
    const int32_t n = m->addr->nCells;
    or_sumA(m, tmp);
    const double xRef = or_average(n, psi);
    for (int32_t c = 0; c < n; c++) tmp[c] *= xRef;
    double s = 0.0;
    for (int32_t c = 0; c < n; c++)
        s += fabs(Apsi[c] - tmp[c]) + fabs(source[c] - tmp[c]);
    return s + OR_SMALL_;
}

/* lduMatrixOperations.C:55-74 */
void or_negSumDiag(const or_addr* a, const double* lower, const double* upper,
                   double* diag)
{
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        diag[a->lower[f]] -= lower[f];
        diag[a->upper[f]] -= upper[f];
    }
}

/* ------------------------------------------------------------------ */
/* preconditioners                                                     */
/* ------------------------------------------------------------------ */

/* DICPreconditioner.C:57-84 */
void or_calcReciprocalD_DIC(const or_matrix* m, double* rD)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++) rD[c] = m->diag[c];
    for (int32_t f = 0; f < a->nFaces; f++)
        rD[a->upper[f]] -= m->upper[f]*m->upper[f]/rD[a->lower[f]];
    for (int32_t c = 0; c < a->nCells; c++) rD[c] = 1.0/rD[c];
}

/* DILUPreconditioner.C:57-85 */
void or_calcReciprocalD_DILU(const or_matrix* m, double* rD)
{
    const or_addr* a = m->addr;
    for (int32_t c = 0; c < a->nCells; c++) rD[c] = m->diag[c];
    for (int32_t f = 0; f < a->nFaces; f++)
        rD[a->upper[f]] -= m->upper[f]*m->lower[f]/rD[a->lower[f]];
    for (int32_t c = 0; c < a->nCells; c++) rD[c] = 1.0/rD[c];
}

static void calc_rD(const or_matrix* m, int precon, double* rD)
{
    const int32_t n = m->addr->nCells;
    switch (precon)
    {
        case OR_PRECON_DIC:
        case OR_PRECON_FDIC:
            or_calcReciprocalD_DIC(m, rD);
            break;
        case OR_PRECON_DILU:
            or_calcReciprocalD_DILU(m, rD);
            break;
        case OR_PRECON_DIAGONAL:   /* diagonalPreconditioner.C:50-63 */
            for (int32_t c = 0; c < n; c++) rD[c] = 1.0/m->diag[c];
            break;
        default:
            break;
    }
}

void or_precondition(const or_matrix* m, int precon, const double* rD,
                     const double* r, double* w)
{
    const or_addr* a = m->addr;
    const int32_t nC = a->nCells, nF = a->nFaces;
    switch (precon)
    {
        case OR_PRECON_NONE:            /* noPreconditioner.C */
            for (int32_t c = 0; c < nC; c++) w[c] = r[c];
            break;
        case OR_PRECON_DIAGONAL:        /* diagonalPreconditioner.C:68-87 */
            for (int32_t c = 0; c < nC; c++) w[c] = rD[c]*r[c];
            break;
        case OR_PRECON_DIC:             /* DICPreconditioner.C:87-123 */
            for (int32_t c = 0; c < nC; c++) w[c] = rD[c]*r[c];
            for (int32_t f = 0; f < nF; f++)
                w[a->upper[f]] -= rD[a->upper[f]]*m->upper[f]*w[a->lower[f]];
            for (int32_t f = nF - 1; f >= 0; f--)
                w[a->lower[f]] -= rD[a->lower[f]]*m->upper[f]*w[a->upper[f]];
            break;
        case OR_PRECON_FDIC:            /* FDICPreconditioner.C:86-125 */
        {
            /* rDuUpper / rDlUpper are precomputed products (:76-81) */
            for (int32_t c = 0; c < nC; c++) w[c] = rD[c]*r[c];
            for (int32_t f = 0; f < nF; f++)
            {
                const double rDuUpper = rD[a->upper[f]]*m->upper[f];
                w[a->upper[f]] -= rDuUpper*w[a->lower[f]];
            }
            for (int32_t f = nF - 1; f >= 0; f--)
            {
                const double rDlUpper = rD[a->lower[f]]*m->upper[f];
                w[a->lower[f]] -= rDlUpper*w[a->upper[f]];
            }
            break;
        }
        case OR_PRECON_DILU:            /* DILUPreconditioner.C:88-135 */
            for (int32_t c = 0; c < nC; c++) w[c] = rD[c]*r[c];
            for (int32_t f = 0; f < nF; f++)
            {
                const int32_t s = a->losort[f];
                w[a->upper[s]] -= rD[a->upper[s]]*m->lower[s]*w[a->lower[s]];
            }
            for (int32_t f = nF - 1; f >= 0; f--)
                w[a->lower[f]] -= rD[a->lower[f]]*m->upper[f]*w[a->upper[f]];
            break;
        default:
            break;
    }
}

/* DILUPreconditioner.C:138-185; DIC/diagonal/none are symmetric */
void or_preconditionT(const or_matrix* m, int precon, const double* rD,
                      const double* r, double* w)
{
    const or_addr* a = m->addr;
    const int32_t nC = a->nCells, nF = a->nFaces;
    if (precon != OR_PRECON_DILU)
    {
        or_precondition(m, precon, rD, r, w);
        return;
    }
    for (int32_t c = 0; c < nC; c++) w[c] = rD[c]*r[c];
    for (int32_t f = 0; f < nF; f++)
        w[a->upper[f]] -= rD[a->upper[f]]*m->upper[f]*w[a->lower[f]];
    for (int32_t f = nF - 1; f >= 0; f--)
    {
        const int32_t s = a->losort[f];
        w[a->lower[s]] -= rD[a->lower[s]]*m->lower[s]*w[a->upper[s]];
    }
}

/* ------------------------------------------------------------------ */
/* smoothers                                                           */
/* ------------------------------------------------------------------ */

/* GaussSeidelSmoother.C:66-187 (forward) and
   symGaussSeidelSmoother.C:66-216 (forward + backward) */
static void gauss_seidel(const or_matrix* m, double* psi, const double* source,
                         int32_t nSweeps, int symmetric)
{
    const or_addr* a = m->addr;
    const int32_t nCells = a->nCells;
    const int32_t* uPtr = a->upper;
    const int32_t* own = a->ownerStart;
    double* bPrime = dalloc(nCells);
    for (int32_t sweep = 0; sweep < nSweeps; sweep++)
    {
        memcpy(bPrime, source, sizeof(double)*nCells);
        /* the interfaces as an effective Jacobi term with the negated
           boundary coefficients and psi before the sweep
           (GaussSeidelSmoother.C:117-143) */
        if (m->interfaces) update_interfaces(m, m->interfaces->bouCoeffs, 1, psi, bPrime);
        int32_t fStart, fEnd = own[0];
        for (int32_t c = 0; c < nCells; c++)
        {
            fStart = fEnd;
            fEnd = own[c + 1];
            double psii = bPrime[c];
            for (int32_t f = fStart; f < fEnd; f++)
                psii -= m->upper[f]*psi[uPtr[f]];
            psii /= m->diag[c];
            for (int32_t f = fStart; f < fEnd; f++)
                bPrime[uPtr[f]] -= m->lower[f]*psii;
            psi[c] = psii;
        }
        if (symmetric)
        {
            fStart = own[nCells];
            for (int32_t c = nCells - 1; c >= 0; c--)
            {
                fEnd = fStart;
                fStart = own[c];
                double psii = bPrime[c];
                for (int32_t f = fStart; f < fEnd; f++)
                    psii -= m->upper[f]*psi[uPtr[f]];
                psii /= m->diag[c];
                for (int32_t f = fStart; f < fEnd; f++)
                    bPrime[uPtr[f]] -= m->lower[f]*psii;
                psi[c] = psii;
            }
        }
    }
    free(bPrime);
}

/* DICSmoother.C:67-116, DILUSmoother.C:67-119, FDICSmoother.C */
static void ilu_smooth(const or_matrix* m, int dilu, const double* rD,
                       double* psi, const double* source, int32_t nSweeps)
{
    const or_addr* a = m->addr;
    const int32_t nC = a->nCells, nF = a->nFaces;
    double* rA = dalloc(nC);
    for (int32_t sweep = 0; sweep < nSweeps; sweep++)
    {
        or_residual(m, psi, source, rA);
        for (int32_t c = 0; c < nC; c++) rA[c] *= rD[c];
        const double* lowerCoeffs = dilu ? m->lower : m->upper;
        for (int32_t f = 0; f < nF; f++)
        {
            const int32_t u = a->upper[f];
            rA[u] -= rD[u]*lowerCoeffs[f]*rA[a->lower[f]];
        }
        for (int32_t f = nF - 1; f >= 0; f--)
        {
            const int32_t l = a->lower[f];
            rA[l] -= rD[l]*m->upper[f]*rA[a->upper[f]];
        }
        for (int32_t c = 0; c < nC; c++) psi[c] += rA[c];
    }
    free(rA);
}

void or_smooth(const or_matrix* m, int smoother, const double* rD,
               double* psi, const double* source, int32_t nSweeps)
{
    switch (smoother)
    {
        case OR_SMOOTH_GS:
            gauss_seidel(m, psi, source, nSweeps, 0);
            break;
        case OR_SMOOTH_SYMGS:
            gauss_seidel(m, psi, source, nSweeps, 1);
            break;
        case OR_SMOOTH_DIC:
        case OR_SMOOTH_FDIC:
            ilu_smooth(m, 0, rD, psi, source, nSweeps);
            break;
        case OR_SMOOTH_DILU:
            ilu_smooth(m, 1, rD, psi, source, nSweeps);
            break;
        case OR_SMOOTH_DICGS:   /* DICGaussSeidelSmoother.C:79-89 */
            ilu_smooth(m, 0, rD, psi, source, nSweeps);
            gauss_seidel(m, psi, source, nSweeps, 0);
            break;
        case OR_SMOOTH_DILUGS:  /* DILUGaussSeidelSmoother.C */
            ilu_smooth(m, 1, rD, psi, source, nSweeps);
            gauss_seidel(m, psi, source, nSweeps, 0);
            break;
        default:
            break;
    }
}

static void smoother_rD(const or_matrix* m, int smoother, double* rD)
{
    switch (smoother)
    {
        case OR_SMOOTH_DIC:
        case OR_SMOOTH_FDIC:
        case OR_SMOOTH_DICGS:
            or_calcReciprocalD_DIC(m, rD);
            break;
        case OR_SMOOTH_DILU:
        case OR_SMOOTH_DILUGS:
            or_calcReciprocalD_DILU(m, rD);
            break;
        default:
            break;
    }
}

/* ------------------------------------------------------------------ */
/* controls and convergence: lduMatrixSolver.C:158-164,                */
/* SolverPerformance.C:31-92                                           */
/* ------------------------------------------------------------------ */

void or_controls_default(or_controls* c)
{
    memset(c, 0, sizeof(*c));
    c->solver = OR_SOLVER_PCG;
    c->preconditioner = OR_PRECON_DIC;
    c->smoother = OR_SMOOTH_GS;
    c->maxIter = 1000;
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
}

static int check_convergence(or_perf* p, double tol, double relTol)
{
    if (p->finalResidual < tol
     || (relTol > OR_SMALL_ && p->finalResidual < relTol*p->initialResidual))
        p->converged = 1;
    else
        p->converged = 0;
    return p->converged;
}

static int check_singularity(or_perf* p, double v)
{
    p->singular = (v < OR_VSMALL_);
    return p->singular;
}

/* ------------------------------------------------------------------ */
/* Krylov solvers                                                      */
/* ------------------------------------------------------------------ */

typedef struct
{
    int kind;                /* OR_PRECON_* */
    double* rD;
    /* GAMG preconditioner state */
    struct gamg_state* gamg;
    int32_t nVcycles;
} precon_state;

static int  gamg_setup(const or_matrix* m, const or_controls* c,
                       struct gamg_state** out);
static void gamg_free(struct gamg_state* g);
static void gamg_precondition(struct gamg_state* g, double* w,
                              const double* r, int32_t nVcycles);

static void precon_setup(const or_matrix* m, const or_controls* c,
                         precon_state* p)
{
    p->kind = c->preconditioner;
    p->rD = NULL;
    p->gamg = NULL;
    if (p->kind == OR_PRECON_GAMG)
    {
        /* GAMGPreconditioner is a GAMGSolver built from the preconditioner
           sub-dictionary (lduMatrixPreconditioner.C:72-82) */
        const or_controls* gc = c->precondControls ? c->precondControls : c;
        p->nVcycles = gc->nVcycles;
        gamg_setup(m, gc, &p->gamg);
        return;
    }
    if (p->kind != OR_PRECON_NONE)
    {
        p->rD = dalloc(m->addr->nCells);
        calc_rD(m, p->kind, p->rD);
    }
}

static void precon_apply(const or_matrix* m, const or_controls* c,
                         precon_state* p, const double* r, double* w, int T)
{
    if (p->kind == OR_PRECON_GAMG)
    {
        gamg_precondition(p->gamg, w, r, p->nVcycles);
        return;
    }
    if (T) or_preconditionT(m, p->kind, p->rD, r, w);
    else   or_precondition(m, p->kind, p->rD, r, w);
}

static void precon_free(precon_state* p)
{
    free(p->rD);
    if (p->gamg) gamg_free(p->gamg);
}

#define HIST(i, v) do { if (history) history[(i)] = (v); } while (0)

/* PCG.C:65-193 */
static int pcg(const or_matrix* m, const or_controls* c, double* psi,
               const double* source, or_perf* perf, double* history)
{
    const int32_t n = m->addr->nCells;
    double* pA = dalloc(n);
    double* wA = dalloc(n);
    double* rA = dalloc(n);
    double wArA = OR_GREAT_, wArAold = wArA;

    or_amul(m, psi, wA);
    for (int32_t i = 0; i < n; i++) rA[i] = source[i] - wA[i];
    const double normFactor = or_normFactor(m, psi, source, wA, pA);
    perf->normFactor = normFactor;
    perf->initialResidual = or_sumMag(n, rA)/normFactor;
    perf->finalResidual = perf->initialResidual;
    HIST(0, perf->initialResidual);

    if (c->minIter > 0 || !check_convergence(perf, c->tolerance, c->relTol))
    {
        precon_state P;
        precon_setup(m, c, &P);
        do
        {
            wArAold = wArA;
            precon_apply(m, c, &P, rA, wA, 0);
            wArA = or_sumProd(n, wA, rA);
            if (perf->nIterations == 0)
            {
                for (int32_t i = 0; i < n; i++) pA[i] = wA[i];
            }
            else
            {
                const double beta = wArA/wArAold;
                for (int32_t i = 0; i < n; i++) pA[i] = wA[i] + beta*pA[i];
            }
            or_amul(m, pA, wA);
            const double wApA = or_sumProd(n, wA, pA);
            if (check_singularity(perf, fabs(wApA)/normFactor)) break;
            const double alpha = wArA/wApA;
            for (int32_t i = 0; i < n; i++)
            {
                psi[i] += alpha*pA[i];
                rA[i] -= alpha*wA[i];
            }
            perf->finalResidual = or_sumMag(n, rA)/normFactor;
            HIST(perf->nIterations + 1, perf->finalResidual);
        } while ((++perf->nIterations < c->maxIter
                  && !check_convergence(perf, c->tolerance, c->relTol))
                 || perf->nIterations < c->minIter);
        precon_free(&P);
    }
    free(pA); free(wA); free(rA);
    return 0;
}

/* PBiCGStab.C:68-254 */
static int pbicgstab(const or_matrix* m, const or_controls* c, double* psi,
                     const double* source, or_perf* perf, double* history)
{
    const int32_t n = m->addr->nCells;
    double* pA = dalloc(n);
    double* yA = dalloc(n);
    double* rA = dalloc(n);

    or_amul(m, psi, yA);
    for (int32_t i = 0; i < n; i++) rA[i] = source[i] - yA[i];
    const double normFactor = or_normFactor(m, psi, source, yA, pA);
    perf->normFactor = normFactor;
    perf->initialResidual = or_sumMag(n, rA)/normFactor;
    perf->finalResidual = perf->initialResidual;
    HIST(0, perf->initialResidual);

    if (c->minIter > 0 || !check_convergence(perf, c->tolerance, c->relTol))
    {
        double* AyA = dalloc(n);
        double* sA = dalloc(n);
        double* zA = dalloc(n);
        double* tA = dalloc(n);
        double* rA0 = dalloc(n);
        memcpy(rA0, rA, sizeof(double)*n);
        double rA0rA = 0, alpha = 0, omega = 0;
        precon_state P;
        precon_setup(m, c, &P);
        do
        {
            const double rA0rAold = rA0rA;
            rA0rA = or_sumProd(n, rA0, rA);
            if (check_singularity(perf, fabs(rA0rA))) break;
            if (perf->nIterations == 0)
            {
                for (int32_t i = 0; i < n; i++) pA[i] = rA[i];
            }
            else
            {
                if (check_singularity(perf, fabs(omega))) break;
                const double beta = (rA0rA/rA0rAold)*(alpha/omega);
                for (int32_t i = 0; i < n; i++)
                    pA[i] = rA[i] + beta*(pA[i] - omega*AyA[i]);
            }
            precon_apply(m, c, &P, pA, yA, 0);
            or_amul(m, yA, AyA);
            const double rA0AyA = or_sumProd(n, rA0, AyA);
            alpha = rA0rA/rA0AyA;
            for (int32_t i = 0; i < n; i++) sA[i] = rA[i] - alpha*AyA[i];
            perf->finalResidual = or_sumMag(n, sA)/normFactor;
            if (++perf->nIterations >= c->minIter
             && check_convergence(perf, c->tolerance, c->relTol))
            {
                for (int32_t i = 0; i < n; i++) psi[i] += alpha*yA[i];
                HIST(perf->nIterations, perf->finalResidual);
                break;
            }
            precon_apply(m, c, &P, sA, zA, 0);
            or_amul(m, zA, tA);
            const double tAtA = or_sumSqr(n, tA);
            omega = or_sumProd(n, tA, sA)/tAtA;
            for (int32_t i = 0; i < n; i++)
            {
                psi[i] += alpha*yA[i] + omega*zA[i];
                rA[i] = sA[i] - omega*tA[i];
            }
            perf->finalResidual = or_sumMag(n, rA)/normFactor;
            HIST(perf->nIterations, perf->finalResidual);
        } while ((perf->nIterations < c->maxIter
                  && !check_convergence(perf, c->tolerance, c->relTol))
                 || perf->nIterations < c->minIter);
        precon_free(&P);
        free(AyA); free(sA); free(zA); free(tA); free(rA0);
    }
    free(pA); free(yA); free(rA);
    return 0;
}

/* PBiCG.C:65-222 */
static int pbicg(const or_matrix* m, const or_controls* c, double* psi,
                 const double* source, or_perf* perf, double* history)
{
    const int32_t n = m->addr->nCells;
    double* pA = dalloc(n);
    double* wA = dalloc(n);
    double* rA = dalloc(n);
    double wArT = 0.0, wArTold;

    or_amul(m, psi, wA);
    for (int32_t i = 0; i < n; i++) rA[i] = source[i] - wA[i];
    const double normFactor = or_normFactor(m, psi, source, wA, pA);
    perf->normFactor = normFactor;
    perf->initialResidual = or_sumMag(n, rA)/normFactor;
    perf->finalResidual = perf->initialResidual;
    HIST(0, perf->initialResidual);

    if (c->minIter > 0 || !check_convergence(perf, c->tolerance, c->relTol))
    {
        double* pT = dalloc(n);
        double* wT = dalloc(n);
        double* rT = dalloc(n);
        double* tmp = dalloc(n);
        /* rT = source - T(psi) */
        or_tmul(m, psi, wT);
        for (int32_t i = 0; i < n; i++) rT[i] = source[i] - wT[i];
        precon_state P;
        precon_setup(m, c, &P);
        do
        {
            wArTold = wArT;
            precon_apply(m, c, &P, rA, wA, 0);
            precon_apply(m, c, &P, rT, wT, 1);
            wArT = or_sumProd(n, wA, rT);
            if (perf->nIterations == 0)
            {
                for (int32_t i = 0; i < n; i++) { pA[i] = wA[i]; pT[i] = wT[i]; }
            }
            else
            {
                const double beta = wArT/wArTold;
                for (int32_t i = 0; i < n; i++)
                {
                    pA[i] = wA[i] + beta*pA[i];
                    pT[i] = wT[i] + beta*pT[i];
                }
            }
            or_amul(m, pA, wA);
            or_tmul(m, pT, wT);
            const double wApT = or_sumProd(n, wA, pT);
            if (check_singularity(perf, fabs(wApT)/normFactor)) break;
            const double alpha = wArT/wApT;
            for (int32_t i = 0; i < n; i++)
            {
                psi[i] += alpha*pA[i];
                rA[i] -= alpha*wA[i];
                rT[i] -= alpha*wT[i];
            }
            perf->finalResidual = or_sumMag(n, rA)/normFactor;
            HIST(perf->nIterations + 1, perf->finalResidual);
        } while ((++perf->nIterations < c->maxIter
                  && !check_convergence(perf, c->tolerance, c->relTol))
                 || perf->nIterations < c->minIter);
        precon_free(&P);
        free(pT); free(wT); free(rT); free(tmp);
    }
    free(pA); free(wA); free(rA);
    return 0;
}

/* smoothSolver.C:77-193 */
static int smooth_solver(const or_matrix* m, const or_controls* c, double* psi,
                         const double* source, or_perf* perf, double* history)
{
    const int32_t n = m->addr->nCells;
    double* rD = dalloc(n);
    smoother_rD(m, c->smoother, rD);
    if (c->nSweeps < 0)
    {
        or_smooth(m, c->smoother, rD, psi, source, -c->nSweeps);
        perf->nIterations -= c->nSweeps;
        free(rD);
        return 0;
    }
    double* Apsi = dalloc(n);
    double* tmp = dalloc(n);
    or_amul(m, psi, Apsi);
    const double normFactor = or_normFactor(m, psi, source, Apsi, tmp);
    perf->normFactor = normFactor;
    for (int32_t i = 0; i < n; i++) tmp[i] = fabs(source[i] - Apsi[i]);
    perf->initialResidual = or_sum(n, tmp)/normFactor;
    perf->finalResidual = perf->initialResidual;
    HIST(0, perf->initialResidual);
    if (c->minIter > 0 || !check_convergence(perf, c->tolerance, c->relTol))
    {
        int k = 0;
        do
        {
            or_smooth(m, c->smoother, rD, psi, source, c->nSweeps);
            or_residual(m, psi, source, tmp);
            perf->finalResidual = or_sumMag(n, tmp)/normFactor;
            HIST(++k, perf->finalResidual);
        } while (((perf->nIterations += c->nSweeps) < c->maxIter
                  && !check_convergence(perf, c->tolerance, c->relTol))
                 || perf->nIterations < c->minIter);
    }
    free(Apsi); free(tmp); free(rD);
    return 0;
}

/* diagonalSolver.C:62-81 */
static int diagonal_solver(const or_matrix* m, double* psi,
                           const double* source, or_perf* perf)
{
    for (int32_t i = 0; i < m->addr->nCells; i++)
        psi[i] = source[i]/m->diag[i];
    perf->nIterations = 0;
    perf->converged = 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* GAMG                                                                */
/* ------------------------------------------------------------------ */

static int or_forward_ = 1;   /* pairGAMGAgglomeration.C:33 */

void or_gamg_reset_direction(void) { or_forward_ = 1; }
int  or_gamg_get_direction(void) { return or_forward_; }

/* pairGAMGAgglomerate.C:123-301 */
int32_t or_pair_agglomerate(const or_addr* a, const double* faceWeights,
                            int32_t* coarseCellMap)
{
    const int32_t nFineCells = a->nCells;
    const int32_t nF = a->nFaces;
    const int32_t* upperAddr = a->upper;
    const int32_t* lowerAddr = a->lower;

    int32_t* cellFaces = ialloc(2*nF);
    int32_t* cellFaceOffsets = ialloc(nFineCells + 1);
    {
        int32_t* nNbrs = ialloc(nFineCells);
        for (int32_t f = 0; f < nF; f++) nNbrs[upperAddr[f]]++;
        for (int32_t f = 0; f < nF; f++) nNbrs[lowerAddr[f]]++;
        cellFaceOffsets[0] = 0;
        for (int32_t c = 0; c < nFineCells; c++)
            cellFaceOffsets[c + 1] = cellFaceOffsets[c] + nNbrs[c];
        memset(nNbrs, 0, sizeof(int32_t)*(nFineCells > 0 ? nFineCells : 1));
        for (int32_t f = 0; f < nF; f++)
        {
            const int32_t u = upperAddr[f];
            cellFaces[cellFaceOffsets[u] + nNbrs[u]] = f;
            nNbrs[u]++;
        }
        for (int32_t f = 0; f < nF; f++)
        {
            const int32_t l = lowerAddr[f];
            cellFaces[cellFaceOffsets[l] + nNbrs[l]] = f;
            nNbrs[l]++;
        }
        free(nNbrs);
    }

    for (int32_t c = 0; c < nFineCells; c++) coarseCellMap[c] = -1;
    int32_t nCoarseCells = 0;

    for (int32_t cellfi = 0; cellfi < nFineCells; cellfi++)
    {
        const int32_t celli = or_forward_ ? cellfi : nFineCells - cellfi - 1;
        if (coarseCellMap[celli] < 0)
        {
            int32_t matchFaceNo = -1;
            double maxFaceWeight = -OR_GREAT;
            for (int32_t fo = cellFaceOffsets[celli]; fo < cellFaceOffsets[celli + 1]; fo++)
            {
                const int32_t f = cellFaces[fo];
                if (coarseCellMap[upperAddr[f]] < 0
                 && coarseCellMap[lowerAddr[f]] < 0
                 && faceWeights[f] > maxFaceWeight)
                {
                    matchFaceNo = f;
                    maxFaceWeight = faceWeights[f];
                }
            }
            if (matchFaceNo >= 0)
            {
                coarseCellMap[upperAddr[matchFaceNo]] = nCoarseCells;
                coarseCellMap[lowerAddr[matchFaceNo]] = nCoarseCells;
                nCoarseCells++;
            }
            else
            {
                int32_t clusterMatchFaceNo = -1;
                double clusterMaxFaceCoeff = -OR_GREAT;
                for (int32_t fo = cellFaceOffsets[celli]; fo < cellFaceOffsets[celli + 1]; fo++)
                {
                    const int32_t f = cellFaces[fo];
                    if (faceWeights[f] > clusterMaxFaceCoeff)
                    {
                        clusterMatchFaceNo = f;
                        clusterMaxFaceCoeff = faceWeights[f];
                    }
                }
                if (clusterMatchFaceNo >= 0)
                {
                    const int32_t cu = coarseCellMap[upperAddr[clusterMatchFaceNo]];
                    const int32_t cl = coarseCellMap[lowerAddr[clusterMatchFaceNo]];
                    coarseCellMap[celli] = cu > cl ? cu : cl;
                }
            }
        }
    }

    for (int32_t cellfi = 0; cellfi < nFineCells; cellfi++)
    {
        const int32_t celli = or_forward_ ? cellfi : nFineCells - cellfi - 1;
        if (coarseCellMap[celli] < 0)
        {
            coarseCellMap[celli] = nCoarseCells;
            nCoarseCells++;
        }
    }

    if (!or_forward_)
    {
        nCoarseCells--;
        for (int32_t c = 0; c < nFineCells; c++)
            coarseCellMap[c] = nCoarseCells - coarseCellMap[c];
        nCoarseCells++;
    }

    or_forward_ = !or_forward_;

    free(cellFaces);
    free(cellFaceOffsets);
    return nCoarseCells;
}

/* GAMGAgglomerateLduAddressing.C:32-239: coarse faces in upper-triangular
   order (by coarse owner, then order of discovery), face restriction and
   flip maps. */
static void agglomerate_ldu_addressing(int32_t nFineFaces,
                                       const int32_t* lowerAddr,
                                       const int32_t* upperAddr,
                                       const int32_t* restrictMap,
                                       int32_t nCoarseCells,
                                       int32_t* nCoarseFacesOut,
                                       int32_t** coarseOwnerOut,
                                       int32_t** coarseNeighbourOut,
                                       int32_t* faceRestrictAddr,
                                       uint8_t* faceFlipMap)
{
    int32_t maxNnbrs = 10;
    int32_t* cCellnFaces = ialloc(nCoarseCells);
    int32_t* cCellFaces = ialloc(maxNnbrs*nCoarseCells);
    int32_t* initCoarseNeighb = ialloc(nFineFaces);
    int32_t nCoarseFaces = 0;

    for (int32_t ff = 0; ff < nFineFaces; ff++)
    {
        const int32_t rmUpper = restrictMap[upperAddr[ff]];
        const int32_t rmLower = restrictMap[lowerAddr[ff]];
        if (rmUpper == rmLower)
        {
            faceRestrictAddr[ff] = -(rmUpper + 1);
        }
        else
        {
            int32_t cOwn = rmUpper, cNei = rmLower;
            if (rmUpper > rmLower) { cOwn = rmLower; cNei = rmUpper; }
            int32_t* ccFaces = &cCellFaces[maxNnbrs*cOwn];
            int found = 0;
            int32_t* ccnFaces = &cCellnFaces[cOwn];
            for (int32_t i = 0; i < *ccnFaces; i++)
            {
                if (initCoarseNeighb[ccFaces[i]] == cNei)
                {
                    found = 1;
                    faceRestrictAddr[ff] = ccFaces[i];
                    break;
                }
            }
            if (!found)
            {
                if (*ccnFaces >= maxNnbrs)
                {
                    const int32_t oldMax = maxNnbrs;
                    maxNnbrs *= 2;
                    cCellFaces = (int32_t*)realloc(cCellFaces, sizeof(int32_t)*maxNnbrs*nCoarseCells);
                    for (int32_t i = nCoarseCells - 1; i >= 0; i--)
                    {
                        int32_t* oldp = &cCellFaces[oldMax*i];
                        int32_t* newp = &cCellFaces[maxNnbrs*i];
                        for (int32_t j = cCellnFaces[i] - 1; j >= 0; j--) newp[j] = oldp[j];
                    }
                    ccFaces = &cCellFaces[maxNnbrs*cOwn];
                }
                ccFaces[*ccnFaces] = nCoarseFaces;
                initCoarseNeighb[nCoarseFaces] = cNei;
                faceRestrictAddr[ff] = nCoarseFaces;
                (*ccnFaces)++;
                nCoarseFaces++;
            }
        }
    }

    int32_t* coarseOwner = ialloc(nCoarseFaces);
    int32_t* coarseNeighbour = ialloc(nCoarseFaces);
    int32_t* coarseFaceMap = ialloc(nCoarseFaces);
    int32_t cf = 0;
    for (int32_t cc = 0; cc < nCoarseCells; cc++)
    {
        int32_t* cFaces = &cCellFaces[maxNnbrs*cc];
        for (int32_t i = 0; i < cCellnFaces[cc]; i++)
        {
            coarseOwner[cf] = cc;
            coarseNeighbour[cf] = initCoarseNeighb[cFaces[i]];
            coarseFaceMap[cFaces[i]] = cf;
            cf++;
        }
    }
    for (int32_t ff = 0; ff < nFineFaces; ff++)
        if (faceRestrictAddr[ff] >= 0)
            faceRestrictAddr[ff] = coarseFaceMap[faceRestrictAddr[ff]];

    for (int32_t ff = 0; ff < nFineFaces; ff++)
    {
        faceFlipMap[ff] = 0;
        const int32_t c = faceRestrictAddr[ff];
        if (c >= 0)
        {
            const int32_t rmUpper = restrictMap[upperAddr[ff]];
            const int32_t rmLower = restrictMap[lowerAddr[ff]];
            if (coarseOwner[c] == rmUpper && coarseNeighbour[c] == rmLower)
                faceFlipMap[ff] = 1;
        }
    }

    free(cCellnFaces); free(cCellFaces); free(initCoarseNeighb); free(coarseFaceMap);
    *nCoarseFacesOut = nCoarseFaces;
    *coarseOwnerOut = coarseOwner;
    *coarseNeighbourOut = coarseNeighbour;
}

/* pairGAMGAgglomerate.C:31-118 (agglomerate levels, with mergeLevels and
   combineLevels GAMGAgglomerateLduAddressing.C:493-611),
   continueAgglomerating GAMGAgglomeration.C:205-230,
   algebraicPair weights algebraicPairGAMGAgglomerationion.C:57-64 */
/* faceAreaPairGAMGAgglomeration.C:55-80: vector / scalar divides each
   component; mag = sqrt of the squares summed x, y, z */
void or_face_area_weights(int32_t nFaces, const double* Sf, const double* magSf, double* w)
{
    for (int32_t f = 0; f < nFaces; f++)
    {
        const double s = sqrt(magSf[f]);
        const double x = (Sf[3*f]/s)*1.0, y = (Sf[3*f + 1]/s)*1.01, z = (Sf[3*f + 2]/s)*1.02;
        w[f] = sqrt(x*x + y*y + z*z);
    }
}

int or_gamg_agglomerate(const or_matrix* m, const or_controls* c,
                        or_hierarchy* h)
{
    const int maxLevels = 50;
    const int32_t mergeLevels = c->mergeLevels > 0 ? c->mergeLevels : 1;
    const int32_t minCells = c->nCellsInCoarsestLevel;
    const or_addr* a0 = m->addr;

    h->nLevels = 0;
    h->nCells = ialloc(maxLevels + 1);
    h->nFaces = ialloc(maxLevels + 1);
    h->lower = (int32_t**)calloc(maxLevels + 1, sizeof(int32_t*));
    h->upper = (int32_t**)calloc(maxLevels + 1, sizeof(int32_t*));
    h->restrictMap = (int32_t**)calloc(maxLevels, sizeof(int32_t*));
    h->faceRestrict = (int32_t**)calloc(maxLevels, sizeof(int32_t*));
    h->faceFlip = (uint8_t**)calloc(maxLevels, sizeof(uint8_t*));
    h->nCells[0] = a0->nCells;
    h->nFaces[0] = a0->nFaces;
    h->lower[0] = (int32_t*)a0->lower;   /* borrowed */
    h->upper[0] = (int32_t*)a0->upper;

    /* level-0 weights: the face areas' (faceAreaPair) or the coefficients'
       (algebraicPairGAMGAgglomeration.C:57-64) */
    double* weights = dalloc(a0->nFaces);
    for (int32_t f = 0; f < a0->nFaces; f++)
    {
        if (c->faceWeights) { weights[f] = c->faceWeights[f]; continue; }
        const double u = fabs(m->upper[f]);
        const double l = fabs(m->lower[f]);
        weights[f] = (m->lower == m->upper) ? u : (u > l ? u : l);
    }

    int32_t nPairLevels = 0, nCreatedLevels = 0;
    while (nCreatedLevels < maxLevels - 1)
    {
        const int32_t lev = nCreatedLevels;
        const int32_t nFine = h->nCells[lev];
        const int32_t nFineFaces = h->nFaces[lev];
        or_addr fa;
        or_addr_init(&fa, nFine, nFineFaces, h->lower[lev], h->upper[lev]);
        int32_t* rmap = ialloc(nFine);
        const int32_t nCoarse = or_pair_agglomerate(&fa, weights, rmap);
        or_addr_free(&fa);

        if (!(nCoarse >= minCells && nCoarse < nFine))
        {
            free(rmap);
            break;
        }

        int32_t* frestrict = ialloc(nFineFaces);
        uint8_t* fflip = (uint8_t*)calloc(nFineFaces > 0 ? nFineFaces : 1, 1);
        int32_t nCoarseFaces;
        int32_t *cOwn, *cNei;
        agglomerate_ldu_addressing(nFineFaces, h->lower[lev], h->upper[lev],
                                   rmap, nCoarse, &nCoarseFaces, &cOwn, &cNei,
                                   frestrict, fflip);

        /* restrictFaceField of the weights (GAMGAgglomerationTemplates.C:137-167) */
        double* cw = dalloc(nCoarseFaces);
        for (int32_t ff = 0; ff < nFineFaces; ff++)
            if (frestrict[ff] >= 0) cw[frestrict[ff]] += weights[ff];
        free(weights);
        weights = cw;

        if (nPairLevels % mergeLevels)
        {
            /* combineLevels(lev): fold level lev into lev-1 */
            const int32_t prev = lev - 1;
            int32_t* prevF = h->faceRestrict[prev];
            uint8_t* prevFlip = h->faceFlip[prev];
            for (int32_t i = 0; i < h->nFaces[prev]; i++)
            {
                if (prevF[i] >= 0)
                {
                    const int32_t fineFacei = prevF[i];
                    prevF[i] = frestrict[fineFacei];
                    prevFlip[i] = fflip[fineFacei];
                }
                else
                {
                    const int32_t fineFacei = -prevF[i] - 1;
                    prevF[i] = -rmap[fineFacei] - 1;
                    prevFlip[i] = fflip[fineFacei];
                }
            }
            int32_t* prevR = h->restrictMap[prev];
            for (int32_t i = 0; i < h->nCells[prev]; i++) prevR[i] = rmap[prevR[i]];
            free(frestrict); free(fflip); free(rmap);
            /* replace mesh level lev (== prev+1) by the new coarse mesh */
            free(h->lower[lev]); free(h->upper[lev]);
            h->lower[lev] = cOwn;
            h->upper[lev] = cNei;
            h->nCells[lev] = nCoarse;
            h->nFaces[lev] = nCoarseFaces;
        }
        else
        {
            h->restrictMap[lev] = rmap;
            h->faceRestrict[lev] = frestrict;
            h->faceFlip[lev] = fflip;
            h->lower[lev + 1] = cOwn;
            h->upper[lev + 1] = cNei;
            h->nCells[lev + 1] = nCoarse;
            h->nFaces[lev + 1] = nCoarseFaces;
            nCreatedLevels++;
        }
        nPairLevels++;
    }
    free(weights);
    h->nLevels = nCreatedLevels;
    return 0;
}

void or_hierarchy_free(or_hierarchy* h)
{
    for (int32_t l = 1; l <= h->nLevels; l++) { free(h->lower[l]); free(h->upper[l]); }
    for (int32_t l = 0; l < h->nLevels; l++)
    {
        free(h->restrictMap[l]); free(h->faceRestrict[l]); free(h->faceFlip[l]);
    }
    free(h->nCells); free(h->nFaces); free(h->lower); free(h->upper);
    free(h->restrictMap); free(h->faceRestrict); free(h->faceFlip);
    memset(h, 0, sizeof(*h));
}

/* GAMGSolverAgglomerateMatrix.C:33-193 (no interfaces) */
void or_agglomerate_matrix(int32_t nFineFaces, int32_t nCoarseCells,
                           int32_t nCoarseFaces, const int32_t* restrictMap,
                           int32_t nFineCells,
                           const int32_t* faceRestrict, const uint8_t* faceFlip,
                           const double* fineDiag, const double* fineUpper,
                           const double* fineLower,
                           double* coarseDiag, double* coarseUpper,
                           double* coarseLower)
{
    for (int32_t c = 0; c < nCoarseCells; c++) coarseDiag[c] = 0.0;
    for (int32_t i = 0; i < nFineCells; i++) coarseDiag[restrictMap[i]] += fineDiag[i];
    for (int32_t f = 0; f < nCoarseFaces; f++) coarseUpper[f] = 0.0;
    if (fineLower)
    {
        for (int32_t f = 0; f < nCoarseFaces; f++) coarseLower[f] = 0.0;
        for (int32_t ff = 0; ff < nFineFaces; ff++)
        {
            const int32_t cf = faceRestrict[ff];
            if (cf >= 0)
            {
                if (!faceFlip[ff])
                {
                    coarseUpper[cf] += fineUpper[ff];
                    coarseLower[cf] += fineLower[ff];
                }
                else
                {
                    coarseUpper[cf] += fineLower[ff];
                    coarseLower[cf] += fineUpper[ff];
                }
            }
            else
            {
                coarseDiag[-1 - cf] += fineUpper[ff] + fineLower[ff];
            }
        }
    }
    else
    {
        for (int32_t ff = 0; ff < nFineFaces; ff++)
        {
            const int32_t cf = faceRestrict[ff];
            if (cf >= 0) coarseUpper[cf] += fineUpper[ff];
            else coarseDiag[-1 - cf] += 2*fineUpper[ff];
        }
    }
}

/* scalarMatrices.C:77-180 (Crout, implicit partial pivoting), row-major */
int or_LUDecompose(int32_t m, double* A, int32_t* pivotIndices)
{
    double* vv = dalloc(m);
    for (int32_t i = 0; i < m; i++)
    {
        double largestCoeff = 0.0, temp;
        const double* Ai = A + (size_t)i*m;
        for (int32_t j = 0; j < m; j++)
            if ((temp = fabs(Ai[j])) > largestCoeff) largestCoeff = temp;
        if (largestCoeff == 0.0) { free(vv); return -1; }
        vv[i] = 1.0/largestCoeff;
    }
    for (int32_t j = 0; j < m; j++)
    {
        double* Aj = A + (size_t)j*m;
        for (int32_t i = 0; i < j; i++)
        {
            double* Ai = A + (size_t)i*m;
            double sum = Ai[j];
            for (int32_t k = 0; k < i; k++) sum -= Ai[k]*A[(size_t)k*m + j];
            Ai[j] = sum;
        }
        int32_t iMax = 0;
        double largestCoeff = 0.0;
        for (int32_t i = j; i < m; i++)
        {
            double* Ai = A + (size_t)i*m;
            double sum = Ai[j];
            for (int32_t k = 0; k < j; k++) sum -= Ai[k]*A[(size_t)k*m + j];
            Ai[j] = sum;
            double temp;
            if ((temp = vv[i]*fabs(sum)) >= largestCoeff)
            {
                largestCoeff = temp;
                iMax = i;
            }
        }
        pivotIndices[j] = iMax;
        if (j != iMax)
        {
            double* AiMax = A + (size_t)iMax*m;
            for (int32_t k = 0; k < m; k++)
            {
                const double t = Aj[k]; Aj[k] = AiMax[k]; AiMax[k] = t;
            }
            vv[iMax] = vv[j];
        }
        if (Aj[j] == 0.0) Aj[j] = OR_SMALL;
        if (j != m - 1)
        {
            const double rDiag = 1.0/Aj[j];
            for (int32_t i = j + 1; i < m; i++) A[(size_t)i*m + j] *= rDiag;
        }
    }
    free(vv);
    return 0;
}

/* scalarMatricesTemplates.C:117-162 */
void or_LUBacksubstitute(int32_t m, const double* LU, const int32_t* pivot,
                         double* x)
{
    int32_t ii = 0;
    for (int32_t i = 0; i < m; i++)
    {
        const int32_t ip = pivot[i];
        double sum = x[ip];
        x[ip] = x[i];
        const double* Li = LU + (size_t)i*m;
        if (ii != 0)
        {
            for (int32_t j = ii - 1; j < i; j++) sum -= Li[j]*x[j];
        }
        else if (sum != 0.0)
        {
            ii = i + 1;
        }
        x[i] = sum;
    }
    for (int32_t i = m - 1; i >= 0; i--)
    {
        double sum = x[i];
        const double* Li = LU + (size_t)i*m;
        for (int32_t j = i + 1; j < m; j++) sum -= Li[j]*x[j];
        x[i] = sum/Li[i];
    }
}

/* GAMG solver state */
struct gamg_level
{
    or_addr addr;
    double *diag, *upper, *lower;   /* lower == upper if symmetric */
    or_matrix m;
    double* rD;                     /* smoother rD */
};

struct gamg_state
{
    const or_matrix* fine;
    or_controls c;
    or_hierarchy h;
    int32_t nLevels;                /* coarse levels */
    struct gamg_level* lev;         /* [nLevels] coarse matrices */
    double* rD0;                    /* finest smoother rD */
    int scale;
    /* coarsest direct solve */
    double* LU;
    int32_t* pivot;
    /* work fields */
    double** coarseCorr;
    double** coarseSrc;
    double *scratch1, *scratch2;
    double *Apsi, *finestCorr, *finestRes;
};

static void gamg_free(struct gamg_state* g)
{
    for (int32_t l = 0; l < g->nLevels; l++)
    {
        or_addr_free(&g->lev[l].addr);
        free(g->lev[l].diag); free(g->lev[l].upper);
        if (g->lev[l].lower != g->lev[l].upper) free(g->lev[l].lower);
        free(g->lev[l].rD);
        free(g->coarseCorr[l]); free(g->coarseSrc[l]);
    }
    free(g->lev); free(g->coarseCorr); free(g->coarseSrc);
    free(g->rD0); free(g->LU); free(g->pivot);
    free(g->scratch1); free(g->scratch2);
    free(g->Apsi); free(g->finestCorr); free(g->finestRes);
    or_hierarchy_free(&g->h);
    free(g);
}

/* GAMGSolver.C:48-332 (ctor: agglomeration, matrix levels, coarsest LU) */
static int gamg_setup(const or_matrix* m, const or_controls* c,
                      struct gamg_state** out)
{
    struct gamg_state* g = (struct gamg_state*)calloc(1, sizeof(*g));
    g->fine = m;
    g->c = *c;
    const int sym = (m->lower == m->upper);
    g->scale = c->scaleCorrection < 0 ? sym : c->scaleCorrection;
    or_gamg_agglomerate(m, c, &g->h);
    g->nLevels = g->h.nLevels;
    const int32_t nL = g->nLevels;
    g->lev = (struct gamg_level*)calloc(nL > 0 ? nL : 1, sizeof(struct gamg_level));
    g->coarseCorr = (double**)calloc(nL > 0 ? nL : 1, sizeof(double*));
    g->coarseSrc = (double**)calloc(nL > 0 ? nL : 1, sizeof(double*));
    int32_t maxSize = m->addr->nCells;
    for (int32_t l = 0; l < nL; l++)
    {
        const double *fd, *fu, *fl;
        if (l == 0) { fd = m->diag; fu = m->upper; fl = sym ? NULL : m->lower; }
        else { fd = g->lev[l-1].diag; fu = g->lev[l-1].upper; fl = sym ? NULL : g->lev[l-1].lower; }
        struct gamg_level* L = &g->lev[l];
        const int32_t nc = g->h.nCells[l + 1], nf = g->h.nFaces[l + 1];
        L->diag = dalloc(nc);
        L->upper = dalloc(nf);
        L->lower = sym ? L->upper : dalloc(nf);
        or_agglomerate_matrix(g->h.nFaces[l], nc, nf, g->h.restrictMap[l],
                              g->h.nCells[l], g->h.faceRestrict[l],
                              g->h.faceFlip[l], fd, fu, fl,
                              L->diag, L->upper, sym ? NULL : L->lower);
        or_addr_init(&L->addr, nc, nf, g->h.lower[l + 1], g->h.upper[l + 1]);
        L->m.addr = &L->addr;
        L->m.diag = L->diag; L->m.upper = L->upper; L->m.lower = L->lower;
        L->rD = dalloc(nc);
        smoother_rD(&L->m, c->smoother, L->rD);
        g->coarseCorr[l] = dalloc(nc);
        g->coarseSrc[l] = dalloc(nc);
        if (nc > maxSize) maxSize = nc;
    }
    g->rD0 = dalloc(m->addr->nCells);
    smoother_rD(m, c->smoother, g->rD0);
    if (nL > 0 && c->directSolveCoarsest)
    {
        /* LUscalarMatrix.C:66-200 + convert(): dense copy, then LUDecompose */
        struct gamg_level* L = &g->lev[nL - 1];
        const int32_t n = L->addr.nCells;
        g->LU = dalloc(n*n);
        g->pivot = ialloc(n);
        for (int32_t i = 0; i < n; i++) g->LU[(size_t)i*n + i] = L->diag[i];
        for (int32_t f = 0; f < L->addr.nFaces; f++)
        {
            const int32_t l = L->addr.lower[f], u = L->addr.upper[f];
            g->LU[(size_t)l*n + u] = L->upper[f];
            g->LU[(size_t)u*n + l] = L->lower[f];
        }
        or_LUDecompose(n, g->LU, g->pivot);
    }
    g->scratch1 = dalloc(maxSize);
    g->scratch2 = dalloc(maxSize);
    g->Apsi = dalloc(m->addr->nCells);
    g->finestCorr = dalloc(m->addr->nCells);
    g->finestRes = dalloc(m->addr->nCells);
    *out = g;
    return 0;
}

/* GAMGSolverScale.C:31-76 */
static void gamg_scale(double* field, double* Acf, const or_matrix* A,
                       const double* source)
{
    const int32_t n = A->addr->nCells;
    or_amul(A, field, Acf);
    double num = 0.0, den = 0.0;
    for (int32_t i = 0; i < n; i++)
    {
        num += source[i]*field[i];
        den += Acf[i]*field[i];
    }
    const double st = den >= 0 ? den + DBL_MIN : den - DBL_MIN;
    const double sf = num/st;
    for (int32_t i = 0; i < n; i++)
        field[i] = sf*field[i] + (source[i] - sf*Acf[i])/A->diag[i];
}

static void restrict_field(double* cf, int32_t nc, const double* ff,
                           int32_t nf, const int32_t* map)
{
    for (int32_t i = 0; i < nc; i++) cf[i] = 0.0;
    for (int32_t i = 0; i < nf; i++) cf[map[i]] += ff[i];
}

static void prolong_field(double* ff, int32_t nf, const double* cf,
                          const int32_t* map)
{
    for (int32_t i = 0; i < nf; i++) ff[i] = cf[map[i]];
}

static int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

/* GAMGSolverSolve.C:520-552 */
static void gamg_solve_coarsest(struct gamg_state* g, double* corr,
                                const double* src)
{
    struct gamg_level* L = &g->lev[g->nLevels - 1];
    const int32_t n = L->addr.nCells;
    if (g->c.directSolveCoarsest)
    {
        memcpy(corr, src, sizeof(double)*n);
        or_LUBacksubstitute(n, g->LU, g->pivot, corr);
    }
    else
    {
        /* GAMGSolver.C:270-322: PCG+DIC (symmetric) or PBiCGStab+DILU,
           with the parent tolerance and relTol, default maxIter/minIter */
        for (int32_t i = 0; i < n; i++) corr[i] = 0.0;
        or_controls cc;
        or_controls_default(&cc);
        const int sym = (L->lower == L->upper);
        cc.solver = sym ? OR_SOLVER_PCG : OR_SOLVER_PBICGSTAB;
        cc.preconditioner = sym ? OR_PRECON_DIC : OR_PRECON_DILU;
        cc.tolerance = g->c.tolerance;
        cc.relTol = g->c.relTol;
        or_perf p;
        or_solve(&L->m, &cc, corr, src, &p, NULL);
    }
}

/* GAMGSolverSolve.C:148-443 */
static void gamg_vcycle(struct gamg_state* g, double* psi, const double* source,
                        double* Apsi, double* finestCorrection,
                        double* finestResidual)
{
    const or_controls* c = &g->c;
    const int32_t coarsest = g->nLevels - 1;
    const or_matrix* fine = g->fine;

    restrict_field(g->coarseSrc[0], g->h.nCells[1], finestResidual,
                   g->h.nCells[0], g->h.restrictMap[0]);

    for (int32_t l = 0; l < coarsest; l++)
    {
        if (c->nPreSweeps)
        {
            const int32_t n = g->h.nCells[l + 1];
            for (int32_t i = 0; i < n; i++) g->coarseCorr[l][i] = 0.0;
            or_smooth(&g->lev[l].m, c->smoother, g->lev[l].rD,
                      g->coarseCorr[l], g->coarseSrc[l],
                      imin(c->nPreSweeps + c->preSweepsLevelMultiplier*l, c->maxPreSweeps));
            if (g->scale && l < coarsest - 1)
                gamg_scale(g->coarseCorr[l], g->scratch1, &g->lev[l].m, g->coarseSrc[l]);
            or_amul(&g->lev[l].m, g->coarseCorr[l], g->scratch1);
            for (int32_t i = 0; i < n; i++) g->coarseSrc[l][i] -= g->scratch1[i];
        }
        restrict_field(g->coarseSrc[l + 1], g->h.nCells[l + 2], g->coarseSrc[l],
                       g->h.nCells[l + 1], g->h.restrictMap[l + 1]);
    }

    gamg_solve_coarsest(g, g->coarseCorr[coarsest], g->coarseSrc[coarsest]);

    for (int32_t l = coarsest - 1; l >= 0; l--)
    {
        const int32_t n = g->h.nCells[l + 1];
        if (c->nPreSweeps) memcpy(g->scratch2, g->coarseCorr[l], sizeof(double)*n);
        prolong_field(g->coarseCorr[l], n, g->coarseCorr[l + 1], g->h.restrictMap[l + 1]);
        if (g->scale && l < coarsest - 1)
            gamg_scale(g->coarseCorr[l], g->scratch1, &g->lev[l].m, g->coarseSrc[l]);
        if (c->nPreSweeps)
            for (int32_t i = 0; i < n; i++) g->coarseCorr[l][i] += g->scratch2[i];
        or_smooth(&g->lev[l].m, c->smoother, g->lev[l].rD,
                  g->coarseCorr[l], g->coarseSrc[l],
                  imin(c->nPostSweeps + c->postSweepsLevelMultiplier*l, c->maxPostSweeps));
    }

    prolong_field(finestCorrection, g->h.nCells[0], g->coarseCorr[0], g->h.restrictMap[0]);
    if (g->scale)
        gamg_scale(finestCorrection, Apsi, fine, finestResidual);
    for (int32_t i = 0; i < g->h.nCells[0]; i++) psi[i] += finestCorrection[i];
    or_smooth(fine, c->smoother, g->rD0, psi, source, c->nFinestSweeps);
}

/* GAMGSolverSolve.C:31-145 */
static int gamg_solver(const or_matrix* m, const or_controls* c, double* psi,
                       const double* source, or_perf* perf, double* history)
{
    const int32_t n = m->addr->nCells;
    double* Apsi = dalloc(n);
    double* finestCorrection = dalloc(n);
    double* finestResidual = dalloc(n);
    or_amul(m, psi, Apsi);
    const double normFactor = or_normFactor(m, psi, source, Apsi, finestCorrection);
    perf->normFactor = normFactor;
    for (int32_t i = 0; i < n; i++) finestResidual[i] = source[i] - Apsi[i];
    perf->initialResidual = or_sumMag(n, finestResidual)/normFactor;
    perf->finalResidual = perf->initialResidual;
    HIST(0, perf->initialResidual);
    if (c->minIter > 0 || !check_convergence(perf, c->tolerance, c->relTol))
    {
        struct gamg_state* g;
        gamg_setup(m, c, &g);
        if (g->nLevels == 0)
        {
            gamg_free(g);
            free(Apsi); free(finestCorrection); free(finestResidual);
            return -2;   /* "No coarse levels created" (GAMGSolver.C:323-331) */
        }
        do
        {
            gamg_vcycle(g, psi, source, Apsi, finestCorrection, finestResidual);
            or_amul(m, psi, Apsi);
            for (int32_t i = 0; i < n; i++) finestResidual[i] = source[i] - Apsi[i];
            perf->finalResidual = or_sumMag(n, finestResidual)/normFactor;
            HIST(perf->nIterations + 1, perf->finalResidual);
        } while ((++perf->nIterations < c->maxIter
                  && !check_convergence(perf, c->tolerance, c->relTol))
                 || perf->nIterations < c->minIter);
        gamg_free(g);
    }
    free(Apsi); free(finestCorrection); free(finestResidual);
    return 0;
}

/* GAMGPreconditioner.C:87-148: nVcycles V-cycles from wA = 0 */
static void gamg_precondition(struct gamg_state* g, double* wA,
                              const double* rA_ss, int32_t nVcycles)
{
    const or_matrix* m = g->fine;
    const int32_t n = m->addr->nCells;
    for (int32_t i = 0; i < n; i++) wA[i] = 0.0;
    double* AwA = g->Apsi;
    double* finestCorrection = g->finestCorr;
    double* finestResidual = g->finestRes;
    memcpy(finestResidual, rA_ss, sizeof(double)*n);
    for (int32_t cycle = 0; cycle < nVcycles; cycle++)
    {
        gamg_vcycle(g, wA, rA_ss, AwA, finestCorrection, finestResidual);
        if (cycle < nVcycles - 1)
        {
            or_amul(m, wA, AwA);
            for (int32_t i = 0; i < n; i++) finestResidual[i] = rA_ss[i] - AwA[i];
        }
    }
}

int or_solve(const or_matrix* m, const or_controls* c, double* psi,
             const double* source, or_perf* perf, double* history)
{
    memset(perf, 0, sizeof(*perf));
    /* lduMatrixSolver.C:50-60: a matrix without off-diagonal coefficients
       (upper == NULL) goes to diagonalSolver */
    if (m->upper == NULL)
        return diagonal_solver(m, psi, source, perf);
    switch (c->solver)
    {
        case OR_SOLVER_PCG:       return pcg(m, c, psi, source, perf, history);
        case OR_SOLVER_PBICGSTAB: return pbicgstab(m, c, psi, source, perf, history);
        case OR_SOLVER_PBICG:     return pbicg(m, c, psi, source, perf, history);
        case OR_SOLVER_SMOOTH:    return smooth_solver(m, c, psi, source, perf, history);
        case OR_SOLVER_GAMG:
            /* coupled interfaces need GAMGInterface agglomeration: not
               restated */
            if (m->interfaces) return -2;
            return gamg_solver(m, c, psi, source, perf, history);
        case OR_SOLVER_DIAGONAL:  return diagonal_solver(m, psi, source, perf);
        default: return -1;
    }
}

/* ------------------------------------------------------------------ */
/* finite-volume assembly                                              */
/* ------------------------------------------------------------------ */

void or_laplacian_assemble(const or_addr* a, const double* gammaCell,
                           const double* weights, const double* magSf,
                           const double* nonOrthDeltaCoeffs,
                           int32_t nB, const int32_t* bFaceCells,
                           const double* bMagSf, const double* bDeltaCoeffs,
                           const int32_t* bType, const double* bValue,
                           double* upper, double* diag,
                           double* internalCoeffs, double* boundaryCoeffs)
{
    /* gammaMagSf = interpolate(gamma)*magSf; upper = deltaCoeffs*gammaMagSf
       (gaussLaplacianSchemes.C:43-56, gaussLaplacianScheme.C:66) */
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        const double gP = gammaCell[a->lower[f]], gN = gammaCell[a->upper[f]];
        const double gf = weights[f]*(gP - gN) + gN;
        const double gammaMagSf = gf*magSf[f];
        upper[f] = nonOrthDeltaCoeffs[f]*gammaMagSf;
    }
    for (int32_t c = 0; c < a->nCells; c++) diag[c] = 0.0;
    or_negSumDiag(a, upper, upper, diag);
    /* gaussLaplacianScheme.C:69-86, non-coupled patches; the boundary gamma
       is the zeroGradient value of the adjacent cell */
    for (int32_t b = 0; b < nB; b++)
    {
        const double pGamma = gammaCell[bFaceCells[b]]*bMagSf[b];
        if (bType[b] == 1)
        {
            internalCoeffs[b] = pGamma*(-bDeltaCoeffs[b]);
            boundaryCoeffs[b] = -pGamma*(bDeltaCoeffs[b]*bValue[b]);
        }
        else
        {
            internalCoeffs[b] = pGamma*0.0;
            boundaryCoeffs[b] = -pGamma*0.0;
        }
    }
}

void or_upwind_assemble(const or_addr* a, const double* phi,
                        int32_t nB, const int32_t* bType, const double* bPhi,
                        const double* bValue,
                        double* lower, double* upper, double* diag,
                        double* internalCoeffs, double* boundaryCoeffs)
{
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        const double w = phi[f] >= 0.0 ? 1.0 : 0.0;   /* pos0 */
        lower[f] = -w*phi[f];
        upper[f] = lower[f] + phi[f];
    }
    for (int32_t c = 0; c < a->nCells; c++) diag[c] = 0.0;
    or_negSumDiag(a, lower, upper, diag);
    for (int32_t b = 0; b < nB; b++)
    {
        if (bType[b] == 1)
        {
            /* fixedValue: valueInternalCoeffs 0, valueBoundaryCoeffs value */
            internalCoeffs[b] = bPhi[b]*0.0;
            boundaryCoeffs[b] = -bPhi[b]*bValue[b];
        }
        else
        {
            /* zeroGradient: valueInternalCoeffs 1, valueBoundaryCoeffs 0 */
            internalCoeffs[b] = bPhi[b]*1.0;
            boundaryCoeffs[b] = -bPhi[b]*0.0;
        }
    }
}

void or_add_boundary(int32_t nB, const int32_t* bFaceCells,
                     const double* internalCoeffs, const double* boundaryCoeffs,
                     double* diag, double* source)
{
    if (diag)
        for (int32_t b = 0; b < nB; b++) diag[bFaceCells[b]] += internalCoeffs[b];
    if (source)
        for (int32_t b = 0; b < nB; b++) source[bFaceCells[b]] += boundaryCoeffs[b];
}

void or_grad_gauss_linear(const or_addr* a, const double* phi,
                          const double* weights, const double* Sf,
                          const double* V,
                          int32_t nB, const int32_t* bFaceCells,
                          const double* bSf, const double* bMagSf,
                          const double* bPhi, const double* bSnGrad,
                          double* grad, double* bGrad)
{
    const int32_t nC = a->nCells;
    for (int32_t i = 0; i < 3*nC; i++) grad[i] = 0.0;
    for (int32_t f = 0; f < a->nFaces; f++)
    {
        const double pP = phi[a->lower[f]], pN = phi[a->upper[f]];
        const double sf = weights[f]*(pP - pN) + pN;
        const int32_t o = a->lower[f], n = a->upper[f];
        for (int d = 0; d < 3; d++)
        {
            const double Sfsf = Sf[3*f + d]*sf;
            grad[3*o + d] += Sfsf;
            grad[3*n + d] -= Sfsf;
        }
    }
    for (int32_t b = 0; b < nB; b++)
    {
        const int32_t c = bFaceCells[b];
        for (int d = 0; d < 3; d++) grad[3*c + d] += bSf[3*b + d]*bPhi[b];
    }
    for (int32_t c = 0; c < nC; c++)
        for (int d = 0; d < 3; d++) grad[3*c + d] /= V[c];
    if (bGrad)
    {
        for (int32_t b = 0; b < nB; b++)
        {
            const int32_t c = bFaceCells[b];
            double n[3];
            for (int d = 0; d < 3; d++) n[d] = bSf[3*b + d]/bMagSf[b];
            const double ng = n[0]*grad[3*c] + n[1]*grad[3*c + 1] + n[2]*grad[3*c + 2];
            for (int d = 0; d < 3; d++)
                bGrad[3*b + d] = grad[3*c + d] + n[d]*(bSnGrad[b] - ng);
        }
    }
}


/* ------------------------------------------------------------------ */
/* mesh geometry                                                       */
/* ------------------------------------------------------------------ */

typedef struct { double x, y, z; } or_vec;

static or_vec v_ld(const double* a, int64_t i) { or_vec v = {a[3*i], a[3*i + 1], a[3*i + 2]}; return v; }
static void v_st(double* a, int64_t i, or_vec v) { a[3*i] = v.x; a[3*i + 1] = v.y; a[3*i + 2] = v.z; }
static or_vec v_add(or_vec a, or_vec b) { or_vec v = {a.x + b.x, a.y + b.y, a.z + b.z}; return v; }
static or_vec v_sub(or_vec a, or_vec b) { or_vec v = {a.x - b.x, a.y - b.y, a.z - b.z}; return v; }
static or_vec v_mul(double s, or_vec a) { or_vec v = {s*a.x, s*a.y, s*a.z}; return v; }
static or_vec v_div(or_vec a, double s) { or_vec v = {a.x/s, a.y/s, a.z/s}; return v; }
/* Vector::operator& (VectorI.H:145-148) */
static double v_dot(or_vec a, or_vec b) { return a.x*b.x + a.y*b.y + a.z*b.z; }
/* Vector::operator^ (VectorI.H:152-160) */
static or_vec v_cross(or_vec a, or_vec b)
{
    or_vec v = {a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x};
    return v;
}
/* mag = sqrt(magSqr), magSqr summed x, y, z (VectorSpaceI.H:392-415) */
static double v_mag(or_vec a) { return sqrt(a.x*a.x + a.y*a.y + a.z*a.z); }

#define OR_VSMALL DBL_MIN           /* doubleScalarVSmall */

/* faceTemplates.C:79-149 */
static void face_area_centre(int32_t n, const int32_t* pts, const double* points,
                             or_vec* area, or_vec* centre)
{
    if (n == 3)
    {
        const or_vec p0 = v_ld(points, pts[0]), p1 = v_ld(points, pts[1]);
        const or_vec p2 = v_ld(points, pts[2]);
        *area = v_mul(1.0/2.0, v_cross(v_sub(p1, p0), v_sub(p2, p0)));
        *centre = v_mul(1.0/3.0, v_add(v_add(p0, p1), p2));
        return;
    }
    or_vec pAvg = {0.0, 0.0, 0.0};
    for (int32_t i = 0; i < n; i++) pAvg = v_add(pAvg, v_ld(points, pts[i]));
    pAvg = v_div(pAvg, (double)n);
    or_vec sumA = {0.0, 0.0, 0.0};
    for (int32_t i = 0; i < n; i++)
    {
        const or_vec p = v_ld(points, pts[i]), pNext = v_ld(points, pts[(i + 1) % n]);
        sumA = v_add(sumA, v_cross(v_sub(pNext, p), v_sub(pAvg, p)));
    }
    /* normalised (VectorSpaceI.H:419-426) */
    const double magA = v_mag(sumA);
    const or_vec zero = {0.0, 0.0, 0.0};
    const or_vec sumAHat = magA > 0 ? v_div(sumA, magA) : zero;
    double sumAn = 0.0;
    or_vec sumAnc = {0.0, 0.0, 0.0};
    for (int32_t i = 0; i < n; i++)
    {
        const or_vec p = v_ld(points, pts[i]), pNext = v_ld(points, pts[(i + 1) % n]);
        const or_vec a = v_cross(v_sub(pNext, p), v_sub(pAvg, p));
        const or_vec c = v_add(v_add(p, pNext), pAvg);
        const double an = v_dot(a, sumAHat);
        sumAn += an;
        sumAnc = v_add(sumAnc, v_mul(an, c));
    }
    *area = v_mul(1.0/2.0, sumA);
    *centre = sumAn > OR_VSMALL ? v_div(v_mul(1.0/3.0, sumAnc), sumAn) : pAvg;
}

void or_face_geometry(int32_t nFaces, const int32_t* faceStart, const int32_t* facePoints,
                      const double* points, double* Cf, double* Sf, double* magSf)
{
    const double rootVSmall = sqrt(OR_VSMALL);
    for (int32_t f = 0; f < nFaces; f++)
    {
        or_vec a, c;
        face_area_centre(faceStart[f + 1] - faceStart[f], facePoints + faceStart[f], points,
                         &a, &c);
        v_st(Cf, f, c);
        v_st(Sf, f, a);
        const double m = v_mag(a);
        magSf[f] = m > rootVSmall ? m : rootVSmall;
    }
}

void or_cell_geometry(int32_t nCells, int32_t nInternal, const int32_t* lower,
                      const int32_t* upper, int32_t nB, const int32_t* bFaceCells,
                      const double* Cf, const double* Sf, const double* bCf, const double* bSf,
                      double* C, double* V)
{
    or_vec* cEst = (or_vec*)calloc((size_t)nCells, sizeof(or_vec));
    or_vec* cc = (or_vec*)calloc((size_t)nCells, sizeof(or_vec));
    int32_t* nCellFaces = (int32_t*)calloc((size_t)nCells, sizeof(int32_t));
    for (int32_t c = 0; c < nCells; c++) V[c] = 0.0;
    /* owner faces (internal, then boundary), then neighbour faces */
    for (int32_t f = 0; f < nInternal; f++)
    {
        cEst[lower[f]] = v_add(cEst[lower[f]], v_ld(Cf, f));
        nCellFaces[lower[f]] += 1;
    }
    for (int32_t b = 0; b < nB; b++)
    {
        cEst[bFaceCells[b]] = v_add(cEst[bFaceCells[b]], v_ld(bCf, b));
        nCellFaces[bFaceCells[b]] += 1;
    }
    for (int32_t f = 0; f < nInternal; f++)
    {
        cEst[upper[f]] = v_add(cEst[upper[f]], v_ld(Cf, f));
        nCellFaces[upper[f]] += 1;
    }
    for (int32_t c = 0; c < nCells; c++) cEst[c] = v_div(cEst[c], (double)nCellFaces[c]);
    for (int32_t f = 0; f < nInternal + nB; f++)
    {
        const int internal = f < nInternal;
        const int32_t o = internal ? lower[f] : bFaceCells[f - nInternal];
        const or_vec fc = internal ? v_ld(Cf, f) : v_ld(bCf, f - nInternal);
        const or_vec fa = internal ? v_ld(Sf, f) : v_ld(bSf, f - nInternal);
        const double pyr3Vol = v_dot(fa, v_sub(fc, cEst[o]));
        const or_vec pc = v_add(v_mul(3.0/4.0, fc), v_mul(1.0/4.0, cEst[o]));
        cc[o] = v_add(cc[o], v_mul(pyr3Vol, pc));
        V[o] += pyr3Vol;
    }
    for (int32_t f = 0; f < nInternal; f++)
    {
        const int32_t nb = upper[f];
        const or_vec fc = v_ld(Cf, f), fa = v_ld(Sf, f);
        const double pyr3Vol = v_dot(fa, v_sub(cEst[nb], fc));
        const or_vec pc = v_add(v_mul(3.0/4.0, fc), v_mul(1.0/4.0, cEst[nb]));
        cc[nb] = v_add(cc[nb], v_mul(pyr3Vol, pc));
        V[nb] += pyr3Vol;
    }
    for (int32_t c = 0; c < nCells; c++)
    {
        const or_vec r = fabs(V[c]) > OR_VSMALL ? v_div(cc[c], V[c]) : cEst[c];
        v_st(C, c, r);
        V[c] *= (1.0/3.0);
    }
    free(cEst);
    free(cc);
    free(nCellFaces);
}

void or_surface_interpolation(int32_t nInternal, const int32_t* lower, const int32_t* upper,
                              const double* C, const double* Cf, const double* Sf,
                              const double* magSf, double* weights, double* deltaCoeffs,
                              double* nonOrthDeltaCoeffs, int32_t nB,
                              const int32_t* bFaceCells, const double* bCf, const double* bSf,
                              const double* bMagSf, double* bDeltaCoeffs,
                              double* bNonOrthDeltaCoeffs)
{
    const double vGreat = DBL_MAX/10;
    for (int32_t f = 0; f < nInternal; f++)
    {
        const or_vec sf = v_ld(Sf, f), cf = v_ld(Cf, f);
        const or_vec cO = v_ld(C, lower[f]), cN = v_ld(C, upper[f]);
        /* surfaceInterpolation.C:174-195 */
        const double SfdOwn = fabs(v_dot(sf, v_sub(cf, cO)));
        const double SfdNei = fabs(v_dot(sf, v_sub(cN, cf)));
        const double SfdOwnNei = SfdOwn + SfdNei;
        if (SfdNei/vGreat < SfdOwnNei) weights[f] = SfdNei/SfdOwnNei;
        else
        {
            const double dOwn = v_mag(v_sub(cf, cO)), dNei = v_mag(v_sub(cN, cf));
            weights[f] = dNei/(dOwn + dNei);
        }
        /* surfaceInterpolation.C:240-243 */
        const or_vec delta = v_sub(cN, cO);
        deltaCoeffs[f] = 1.0/v_mag(delta);
        /* surfaceInterpolation.C:300-315 */
        const or_vec unitArea = v_div(sf, magSf[f]);
        const double a = v_dot(unitArea, delta), b = 0.05*v_mag(delta);
        nonOrthDeltaCoeffs[f] = 1.0/(a > b ? a : b);
    }
    for (int32_t k = 0; k < nB; k++)
    {
        /* fvPatch::nf, fvPatch::delta (fvPatch.C:138-164) */
        const or_vec nf = v_div(v_ld(bSf, k), bMagSf[k]);
        const or_vec d = v_sub(v_ld(bCf, k), v_ld(C, bFaceCells[k]));
        const or_vec delta = v_mul(v_dot(nf, d), nf);
        bDeltaCoeffs[k] = 1.0/v_mag(delta);
        const double a = v_dot(nf, delta), b = 0.05*v_mag(delta);
        bNonOrthDeltaCoeffs[k] = 1.0/(a > b ? a : b);
    }
}
