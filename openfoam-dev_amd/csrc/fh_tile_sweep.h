// Tiled wavefront sweeps for hex-box meshes (fh_tile_sweep.hip).
#pragma once

#include "fh_kernels.h"
#include "fh_solver_kernels.h"

#include <hip/hip_fp16.h>

namespace fh
{

// tile decomposition of an nx x ny x nz box into bx x by x bz tiles
struct TileView
{
    int32_t nx, ny, nz;
    int32_t bx, by, bz;
    int32_t nTI, nTJ, nTK;
    const int32_t* order;     // tiles in wavefront order (ascending TI+TJ+TK)
};

struct TileSync
{
    unsigned int* flags;               // per tile: launch epoch + 1 when done
    unsigned long long* ticket;        // monotonic ticket counter
    unsigned long long ticketBase;     // tickets handed out by earlier launches
    unsigned long long epoch;          // launches made with this context
};

// tile depth per kind and precision (LDS budget: the tile's staged inputs)
template<class T> struct TileBZ;
// (alt: the shallower tiles of the global preconditioner sweeps)
template<> struct TileBZ<double> { static constexpr int pre = 8; static constexpr int gs = 4; static constexpr int alt = 4; };
template<> struct TileBZ<float> { static constexpr int pre = 16; static constexpr int gs = 8; static constexpr int alt = 8; };

template<class T>
inline int tile_bz(int kind)
{
    return (kind == SWEEP_GS_FWD || kind == SWEEP_GS_BWD) ? TileBZ<T>::gs : TileBZ<T>::pre;
}

// local: block-Jacobi over tiles (each tile on its own, no inter-tile waits)
template<class T>
void launch_tile_sweep(int kind, const TileView& G, const RowsView& R,
                       const SweepData<T>& d, TileSync& sync, DevState* st, hipStream_t s,
                       bool local = false);

// one-wavefront global DIC / DILU pass over 8 x 8 x 16 tiles (packed
// coefficients d.dir required); backward with d.dotR also sums w*dotR into
// one partial per tile.  Returns the tile count.
template<class T>
int32_t launch_sweep_wave(bool backward, const TileView& G, const SweepData<T>& d,
                          TileSync& sync, DevState* st, hipStream_t s,
                          const FusedFinalize& F = FusedFinalize());

// direction-packed coefficients of a box: dst[0..5] = lower K, J, I and
// upper I, J, K per cell (nDst 6), or the lower three only (nDst 3)
template<class T>
void launch_pack_dir_coefs(const TileView& G, const RowsView& R, const T* val, int nDst,
                           T* const dst[6], hipStream_t s);
// FP32 tile-local DIC / DILU on direction-packed coefficients (mixed inner
// preconditioner)
// forward: a = r, backward: a = w (in place); cK, cJ, cI the coefficients of
// the pass (lower forward, upper backward).  Backward with partials: also
// sum w*r into one partial per workgroup; returns the workgroup count.
// both passes in one launch (nx a multiple of 16; returns -1 otherwise,
// nothing launched): w = (tile-local DIC/DILU)^-1 x from the packed
// coefficients (symmetric: the lower three only), with sum w*x into one
// partial per workgroup when partials is given; returns the partial count
int32_t launch_tile_local_apply_f32(const TileView& G, const float* x, const float* rD,
                                    const float* const dir[6], bool symmetric, float* w,
                                    const DevState* st, hipStream_t s, double* partials,
                                    const __half* const half3[3] = nullptr,
                                    const FusedFinalize& F = FusedFinalize(),
                                    double* tileSum = nullptr, float halfScale = 1.f);
// FP16 copies of three FP32 arrays of n values, scaled by scale (a power of
// two)
void launch_pack_half(int64_t n, const float* const src[3], __half* const dst[3], float scale,
                      hipStream_t s);
// b = a rounded to FP16 precision at the storage scale (kept in FP32,
// unscaled)
void launch_round_half(int64_t n, const float* a, float* b, float scale, hipStream_t s);
// range[0] = bits of max |a|, range[1] = bits of min non-zero |a| (the
// caller initialises them to 0 and 0xffffffff)
void launch_abs_range(int64_t n, const float* a, unsigned int* range, hipStream_t s);
int32_t launch_tile_local_f32(bool backward, const TileView& G, const float* a, const float* rD,
                              const float* cK, const float* cJ, const float* cI, float* w,
                              const DevState* st, hipStream_t s, const float* r = nullptr,
                              double* partials = nullptr);
// a box field (rows of nx cells) copied to rows of nxp cells (padding 0),
// and back
void launch_pad_rows(int32_t nx, int32_t nxp, int64_t rows, const float* src, float* dst,
                     hipStream_t s);
void launch_unpad_rows(int32_t nx, int32_t nxp, int64_t rows, const float* src, float* dst,
                       hipStream_t s);

} // namespace fh
