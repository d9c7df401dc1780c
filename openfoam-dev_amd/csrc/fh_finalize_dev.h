// Device side of the solver finalize: the scalar logic of each FinalizeOp
// on the summed block partials, and the last-block arrival that lets a
// reduction kernel apply it itself (included by the kernel sources only).
#pragma once

#include "fh_solver_kernels.h"

#include <cfloat>
#include <cmath>

namespace fh
{

__device__ inline bool dev_converged(DevState* st)
{
    // SolverPerformance.C:59-92
    const bool c = st->finalResidual < st->tolerance
                || (st->relTol > 1e-20 && st->finalResidual < st->relTol*st->initialResidual);
    st->converged = c ? 1 : 0;
    return c;
}

__device__ inline bool dev_singular(DevState* st, double v)
{
    st->singular = (v < DBL_MIN) ? 1 : 0;   // SolverPerformance.C:31-44
    return st->singular;
}

// The scalar logic of a finalize op on the summed partials (a, b).
__device__ inline void finalize_apply(double a, double b, int32_t op, DevState* st, int32_t n,
                               int32_t seq)
{
    switch (op)
    {
        case FIN_SUM:
            st->sum0 = a;
            st->sum1 = b;
            break;
        case FIN_XREF:     // gAverage(psi) (FieldReductionFunctions.C:334-357)
            st->sum0 = a/n;
            break;
        case FIN_INIT:     // normFactor, initial residual, first convergence test
        {
            st->normFactor = a + 1e-20;
            st->initialResidual = b/st->normFactor;
            st->finalResidual = st->initialResidual;
            st->nIterations = 0;
            st->singular = 0;
            st->converged = 0;
            st->earlySeq = -1;
            if (!isfinite(st->initialResidual)) { st->numericError = 1; st->done = 1; break; }
            if (st->minIter > 0) st->done = 0;
            else st->done = dev_converged(st) ? 1 : 0;
            break;
        }
        case FIN_PCG_WARA:
            st->wArAold = st->wArA;
            st->wArA = a;
            if (st->nIterations > 0) st->beta = st->wArA/st->wArAold;
            break;
        case FIN_PCG_WAPA:
            st->wApA = a;
            if (dev_singular(st, fabs(a)/st->normFactor)) { st->done = 1; break; }
            st->alpha = st->wArA/st->wApA;
            break;
        case FIN_PCG_RES:
        {
            st->finalResidual = a/st->normFactor;
            if (!isfinite(st->finalResidual)) { st->numericError = 1; st->done = 1; break; }
            st->nIterations++;
            bool cont = false;
            if (st->nIterations < st->maxIter) cont = !dev_converged(st);
            cont = cont || st->nIterations < st->minIter;
            if (!cont) st->done = 1;
            break;
        }
        case FIN_BICG_START:
        {
            // first iteration: rA0rA = rA0 & rA
            st->rA0rAold = st->rA0rA;
            st->rA0rA = a;
            if (dev_singular(st, fabs(st->rA0rA))) { st->done = 1; break; }
            break;
        }
        case FIN_BICG_ALPHA:
            st->rA0AyA = a;
            st->bicgAlpha = st->rA0rA/st->rA0AyA;
            break;
        case FIN_BICG_S:
        {
            st->finalResidual = a/st->normFactor;
            if (!isfinite(st->finalResidual)) { st->numericError = 1; st->done = 1; break; }
            st->nIterations++;
            if (st->nIterations >= st->minIter && dev_converged(st))
            {
                st->earlySeq = seq;
                st->done = 1;
            }
            break;
        }
        case FIN_BICG_OMEGA:
            st->tAtA = a;
            st->tAsA = b;
            st->omega = st->tAsA/st->tAtA;
            break;
        case FIN_BICG_RES:
        {
            st->finalResidual = a/st->normFactor;
            if (!isfinite(st->finalResidual)) { st->numericError = 1; st->done = 1; break; }
            bool cont = false;
            if (st->nIterations < st->maxIter) cont = !dev_converged(st);
            cont = cont || st->nIterations < st->minIter;
            if (!cont) { st->done = 1; break; }
            // start of the next iteration (PBiCGStab.C:149-185)
            st->rA0rAold = st->rA0rA;
            st->rA0rA = b;
            if (dev_singular(st, fabs(st->rA0rA))) { st->done = 1; break; }
            if (dev_singular(st, fabs(st->omega))) { st->done = 1; break; }
            st->bicgBeta = (st->rA0rA/st->rA0rAold)*(st->bicgAlpha/st->omega);
            break;
        }
        case FIN_INNER_INIT:
        {
            // inner FP32 solve of the mixed path: relative reduction of
            // sum |r| only (tolerance 0), normFactor 1
            st->normFactor = 1.0;
            st->initialResidual = b;
            st->finalResidual = b;
            st->nIterations = 0;
            st->singular = 0;
            st->converged = 0;
            st->earlySeq = -1;
            st->done = (b > 0.0 && isfinite(b)) ? 0 : 1;
            break;
        }
        case FIN_OUTER_RES:
        {
            st->finalResidual = b/st->normFactor;
            if (!isfinite(st->finalResidual)) { st->numericError = 1; st->done = 1; break; }
            st->done = dev_converged(st) ? 1 : 0;
            break;
        }
        case FIN_MIXED_RES:
        {
            // FP32 recurrence residual; schedule an FP64 replacement when it
            // fell by replaceRatio since the last one, when it looks
            // converged, or at the iteration limit
            const double est = a/st->normFactor;
            st->foldPending = 0;     // this iteration's update folded it
            if (!isfinite(est)) { st->numericError = 1; st->done = 1; break; }
            st->nIterations++;
            bool looks = est < st->tolerance
                      || (st->relTol > 1e-20 && est < st->relTol*st->initialResidual);
            st->replace = (looks || est < st->replaceRatio*st->resLast
                           || st->nIterations >= st->maxIter) ? 1 : 0;
            break;
        }
        case FIN_MIXED_TRUE:
        {
            if (!st->replace) break;
            st->replace = 0;
            st->nReplace++;
            st->foldPending = 1;     // the next update folds the FP32 correction
            st->finalResidual = b/st->normFactor;
            if (!isfinite(st->finalResidual)) { st->numericError = 1; st->done = 1; break; }
            st->resLast = st->finalResidual;
            bool cont = false;
            if (st->nIterations < st->maxIter) cont = !dev_converged(st);
            cont = cont || st->nIterations < st->minIter;
            if (!cont) st->done = 1;
            break;
        }
        case FIN_SMOOTH_RES:
        {
            // smoothSolver.C:156-189: nIterations += nSweeps before the test
            st->finalResidual = b/st->normFactor;
            if (!isfinite(st->finalResidual)) { st->numericError = 1; st->done = 1; break; }
            st->nIterations += seq;
            bool cont = false;
            if (st->nIterations < st->maxIter) cont = !dev_converged(st);
            cont = cont || st->nIterations < st->minIter;
            if (!cont) st->done = 1;
            break;
        }
        default:
            break;
    }
}

__device__ __forceinline__ void fin_store(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double fin_load(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned int fin_arrive(unsigned int* c)
{
    return __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fused finalize.  Every block publishes its partial pair (a, b, valid in
// thread 0) at slot with device-coherent stores and, once they have
// completed, counts itself into its group of 64 slots.  The last block of a
// group sums the group's partials (one lane each, fixed butterfly order)
// into a group pair stored after the 2 nb partials and counts the group in;
// the last group sums the group pairs the same way, applies F.op to F.st and
// resets the counters (F.counter[0]: groups, F.counter[S*(1 + g)]: group g,
// S = FIN_COUNTER_STRIDE) for the next launch.  Every thread of every block of the launch calls it with
// the same nb (a block that leaves early would leave a count short); the
// sums depend on nb only, not on the arrival order.  Without F.counter the
// block only writes its partials and a separate finalize follows.
__device__ inline void partials_arrive(double a, double b, double* partials, int32_t nb,
                                       int32_t slot, const FusedFinalize& F)
{
    __shared__ int lastInGroup;
    if (!F.counter)
    {
        if (threadIdx.x == 0)
        {
            partials[slot] = a;
            partials[nb + slot] = b;
        }
        return;
    }
    const int32_t g = slot >> 6;
    const int32_t nG = (nb + 63) >> 6;
    if (threadIdx.x == 0)
    {
        fin_store(&partials[slot], a);
        fin_store(&partials[nb + slot], b);
        // the partials reach the device coherence point before the count
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const unsigned int inG = (unsigned int)min(64, nb - (g << 6));
        lastInGroup = fin_arrive(&F.counter[FIN_COUNTER_STRIDE*(1 + g)]) == inG - 1;
    }
    __syncthreads();
    if (!lastInGroup || threadIdx.x >= 64) return;
    // wavefront 0 of the group's last block
    const int lane = threadIdx.x;
    const int32_t i = (g << 6) + lane;
    double sa = i < nb ? fin_load(&partials[i]) : 0.0;
    double sb = i < nb ? fin_load(&partials[nb + i]) : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        sa += __shfl_xor(sa, off, 64);
        sb += __shfl_xor(sb, off, 64);
    }
    double* grp = partials + 2*(int64_t)nb;
    int last = 0;
    if (lane == 0)
    {
        fin_store(&grp[g], sa);
        fin_store(&grp[nG + g], sb);
        __hip_atomic_store(&F.counter[FIN_COUNTER_STRIDE*(1 + g)], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        last = fin_arrive(&F.counter[0]) == (unsigned int)(nG - 1);
    }
    last = __shfl(last, 0, 64);
    if (!last) return;
    // the last group: sum the group pairs (lane l takes l, l + 64, ...)
    double ga = 0.0, gb = 0.0;
    for (int32_t q = lane; q < nG; q += 64)
    {
        ga += fin_load(&grp[q]);
        gb += fin_load(&grp[nG + q]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        ga += __shfl_xor(ga, off, 64);
        gb += __shfl_xor(gb, off, 64);
    }
    if (lane == 0)
    {
        finalize_apply(ga, gb, F.op, F.st, F.n, 0);
        __hip_atomic_store(&F.counter[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

} // namespace fh
