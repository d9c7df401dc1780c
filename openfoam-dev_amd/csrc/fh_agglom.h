// Pairwise GAMG agglomeration on the device (fh_agglom.hip).
#pragma once

#include "fh_internal.h"

#include <memory>
#include <vector>

namespace fh
{

// One level of the agglomeration hierarchy in natural numbering, as
// GAMGAgglomeration holds it (GAMGAgglomeration.H: restrictAddressing,
// faceRestrictAddressing, faceFlipMap, and the coarse level's lduAddressing)
struct AggMaps
{
    int32_t nFine = 0, nFineFaces = 0, nCoarse = 0, nCoarseFaces = 0;
    DBuf<int32_t> restrictMap;     // [nFine] fine cell -> coarse cell
    DBuf<int32_t> faceRestrict;    // [nFineFaces] coarse face, or -1 - coarse cell
    DBuf<uint8_t> faceFlip;        // [nFineFaces]
    DBuf<int32_t> lower, upper;    // [nCoarseFaces] coarse face addressing
    DBuf<double> weights;          // [nCoarseFaces] restricted face weights
};

// The pair-agglomeration hierarchy of a mesh (pairGAMGAgglomerate.C:31-301,
// GAMGAgglomerateLduAddressing.C:32-239 and combineLevels :514-611,
// continueAgglomerating GAMGAgglomeration.C:205-230) from the device face
// addressing and face weights; `forward` is the process-global direction
// toggle (pairGAMGAgglomeration::forward_), flipped after every pair level.
std::vector<std::unique_ptr<AggMaps>> device_agglomerate(int32_t nCells, int32_t nFaces,
                                                         const int32_t* dLower,
                                                         const int32_t* dUpper,
                                                         const double* dWeights,
                                                         int32_t minCells, int32_t mergeLevels,
                                                         bool& forward, hipStream_t s);

// The transfer lists of one level in the library row numbering of the fine
// and coarse meshes (perm[row] = natural cell): rStart/rCells the fine rows
// of each coarse row (ascending fine natural order), prolong the coarse row
// of each fine row, iStart/iFaces the fine faces inside each coarse row's
// cell, fStart/fFaces/fFlip the fine faces of each coarse face (coarse face
// order); every list ascending, the order the reference adds in
// (GAMGAgglomerationTemplates.C:76-219, GAMGSolverAgglomerateMatrix.C:33-193).
struct TransferLists
{
    DBuf<int32_t> rStart, rCells, prolong, iStart, iFaces, fStart, fFaces;
    DBuf<uint8_t> fFlip;
};
void device_transfer_lists(const AggMaps& maps, const int32_t* finePerm,
                           const int32_t* coarsePerm, TransferLists& T, hipStream_t s);

// The algebraicPair face weights of a matrix: |upper|, or max(|upper|,
// |lower|) when lowerF is given (algebraicPairGAMGAgglomeration.C:57-64).
void launch_agg_weights(int32_t nF, const double* upperF, const double* lowerF, double* w,
                        hipStream_t s);

// Exclusive prefix sum of n int32 values: out[0..n] with out[n] the total
// (out may not alias in).
void device_exclusive_scan(const int32_t* in, int32_t* out, int32_t n, hipStream_t s);

} // namespace fh
