// Internal declarations of the foamhip library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "foamhip.h"

namespace fh
{

// ---------------------------------------------------------------------------
// errors: internal code throws fh::Error, the C entry points catch it,
// record the message for fh_last_error() and return the status.
// ---------------------------------------------------------------------------
struct Error
{
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...);
// runs fn(ctx), turning fh::Error and allocation failures into a status
int guarded_call(void (*fn)(void*), void* ctx);
void set_last_error(const std::string& msg);

#define FH_HIP(call)                                                         \
    do                                                                       \
    {                                                                        \
        hipError_t fh_e_ = (call);                                           \
        if (fh_e_ != hipSuccess)                                             \
        {                                                                    \
            ::fh::fail(fh_e_ == hipErrorOutOfMemory ? FH_ERR_NOMEM           \
                                                     : FH_ERR_HIP,           \
                       "%s failed: %s (%s:%d)", #call,                       \
                       hipGetErrorString(fh_e_), __FILE__, __LINE__);        \
        }                                                                    \
    } while (0)

// ---------------------------------------------------------------------------
// owning device buffer
// ---------------------------------------------------------------------------
template<class T>
struct DBuf
{
    T* p = nullptr;
    size_t n = 0;

    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }

    void alloc(size_t count)
    {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        FH_HIP(hipMalloc(&p, count*sizeof(T)));
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s)
    {
        alloc(count);
        if (count)
            FH_HIP(hipMemcpyAsync(p, h, count*sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    size_t bytes() const { return n*sizeof(T); }
};

// ---------------------------------------------------------------------------
// sweep schedule: rows renumbered by (part, level, original index); each
// part is walked level by level by one workgroup, waiting on the progress
// counters of the parts it depends on.
// ---------------------------------------------------------------------------
struct Schedule
{
    int32_t nParts = 1;
    int32_t nLevels = 0;
    // per part: [partLevelOff[p], partLevelOff[p+1]) indexes segStart; the
    // part's first level is partMinLevel[p]; rows of level (minL + k) are
    // [segStart[off+k], segStart[off+k+1])
    std::vector<int32_t> partLevelOff;
    std::vector<int32_t> partMinLevel;
    std::vector<int32_t> segStart;
    // parts whose rows are lower (up) / upper (down) neighbours of part p
    std::vector<int32_t> upOff, upList, downOff, downList;
    // ticket order: forward sweeps take parts in order[], backward in reverse
    std::vector<int32_t> order;

    DBuf<int32_t> d_partLevelOff, d_partMinLevel, d_segStart;
    DBuf<int32_t> d_upOff, d_upList, d_downOff, d_downList, d_order;
    int32_t maxSegRows = 0;
};

struct fh_mesh_impl;

} // namespace fh

// Public opaque types
struct fh_mesh
{
    int device = 0;
    int32_t nCells = 0;
    int32_t nFaces = 0;
    std::vector<int32_t> lower, upper;      // caller's face order
    std::vector<int32_t> perm, iperm;       // perm[new] = old cell
    std::vector<int32_t> levelOf;           // per old cell

    // SELL-64 row structure in the new numbering
    int32_t nSlices = 0;
    int64_t nnzPad = 0;
    std::vector<int32_t> sliceStart;        // [nSlices+1]
    std::vector<int32_t> rowInfo;           // nLower | nEnt << 16
    std::vector<int32_t> col;               // SELL layout, new numbering
    std::vector<int32_t> face;              // (f << 1 | isUpper) or -1 pad

    fh::DBuf<int32_t> d_sliceStart, d_rowInfo, d_col, d_face;
    fh::DBuf<int32_t> d_perm;
    fh::DBuf<int32_t> d_lower, d_upper;     // face-ordered addressing

    fh::Schedule sched;
    int32_t boxNx = 0, boxNy = 0, boxNz = 0;

    // hex box in natural numbering: tiled wavefront sweeps (16 x 16 x bz
    // tiles, bz = 4, 8 or 16 by kind and precision)
    bool tiled = false;
    struct TileShape
    {
        int32_t bx = 0, bz = 0, nTI = 0, nTJ = 0, nTK = 0;   // bx x bx x bz tiles
        fh::DBuf<int32_t> d_order;
    };
    TileShape tiles[4];
    int32_t maxTiles = 0;
    hipStream_t setupStream = nullptr;

    // chunks of the sync-free sweep (one-part meshes whose levels are wide
    // enough): rows of one level, at most 64 per chunk
    bool chunkSweep = false;
    int32_t nChunks = 0;
    fh::DBuf<int32_t> d_chunkStart, d_fwdOff, d_fwdDeps, d_bwdOff, d_bwdDeps;

    // GAMG agglomeration cached with the mesh (built on the first GAMG
    // solve, as GAMGAgglomeration is cached on the mesh database)
    std::shared_ptr<void> gamgAgg;
    // faceAreaPair agglomeration weights, one per internal face (empty
    // until the face areas are given)
    std::vector<double> faceAreaWeights;

    // coupled (cyclic) interfaces: the caller's interface faces (concatenated
    // in interface order) regrouped per library row in that order; entry e
    // of row r couples it to ifNbr[e] with the caller's face ifFace[e]
    int32_t nInterfaces = 0;
    int32_t nIfFaces = 0;
    std::vector<int32_t> ifStart, ifNbr, ifFace;   // [nCells+1], [nIfFaces], [nIfFaces]
    fh::DBuf<int32_t> d_ifStart, d_ifNbr, d_ifFace;

    // multi-colour row order (greedy over the caller's cell order; red-black
    // on a box), built on first use by the FP32 Gauss-Seidel of the mixed
    // GAMG V-cycles
    bool coloured = false;
    std::vector<int32_t> colourStart;       // [nColours + 1]
    fh::DBuf<int32_t> d_colourRows;         // library rows grouped by colour
    fh::DBuf<int32_t> d_colourStart;        // device copy of colourStart
    int32_t maxColourRows = 0;              // rows of the largest colour
};

namespace fh
{
// mesh construction (fh_mesh.cpp)
void build_mesh(fh_mesh* m, int32_t nCells, int32_t nFaces,
                const int32_t* lower, const int32_t* upper, bool allowBox = true);
// the per-row interface entries of a coupled mesh (after build_mesh)
void build_interfaces(fh_mesh* m, int32_t nInterfaces, const int32_t* start,
                      const int32_t* faceCells, const int32_t* nbrCells);
size_t mesh_device_bytes(const fh_mesh* m);
} // namespace fh
