// Mesh setup: LDU face addressing -> level-ordered SELL-64 row structure
// and the partitioned sweep schedule.  Done once per mesh on the host.
//
// Row order.  The reference triangular recurrences (DIC/DILU sweeps,
// Gauss-Seidel, the DIC/DILU factor, DICPreconditioner.C:57-123) run over
// faces in upper-triangular order; row u only ever reads rows l < u that
// share a face with it.  Any renumbering in which every such l precedes u
// therefore reproduces them exactly.  Cells are renumbered by
// (part, level, original index), where level(u) = 1 + max level(l) over the
// lower neighbours of u: the rows of one level of one part are contiguous,
// so a workgroup that owns a part reads each level as one coalesced
// segment.  Per-row accumulation keeps the reference face order
// (lduMatrixATmul.C:66-79), so sums are bitwise identical.
#include "fh_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cstdio>

namespace fh
{

static int column_size_for(int32_t nx, int32_t ny)
{
    if (const char* s = std::getenv("FOAMHIP_COLUMN"))
    {
        const int v = std::atoi(s);
        if (v >= 1) return v;
    }
    return (nx >= 192 && ny >= 192) ? 16 : 8;
}

// blockMesh hex box: cells i + nx*(j + ny*k) (blockCreate.C:349-377), faces
// per owner in increasing neighbour order (polyMeshFromShapeMesh.C:170-281).
static bool detect_box(const fh_mesh* m, int32_t& nx, int32_t& ny, int32_t& nz)
{
    nx = ny = nz = 0;
    const int32_t N = m->nCells, F = m->nFaces;
    if (N < 8 || F < 3) return false;
    if (m->lower[0] != 0 || m->lower[1] != 0 || m->lower[2] != 0) return false;
    if (F > 3 && m->lower[3] == 0) return false;
    if (m->upper[0] != 1) return false;
    const int64_t a = m->upper[1], b = m->upper[2];
    if (a < 2 || b % a != 0) return false;
    const int64_t X = a, Y = b/a;
    if (Y < 2 || N % b != 0) return false;
    const int64_t Z = N/b;
    if (Z < 2) return false;
    const int64_t expF = (X - 1)*Y*Z + X*(Y - 1)*Z + X*Y*(Z - 1);
    if (expF != F) return false;
    int64_t f = 0;
    for (int64_t k = 0; k < Z; k++)
        for (int64_t j = 0; j < Y; j++)
            for (int64_t i = 0; i < X; i++)
            {
                const int64_t c = i + X*(j + Y*k);
                if (i < X - 1)
                {
                    if (m->lower[f] != c || m->upper[f] != c + 1) return false;
                    f++;
                }
                if (j < Y - 1)
                {
                    if (m->lower[f] != c || m->upper[f] != c + X) return false;
                    f++;
                }
                if (k < Z - 1)
                {
                    if (m->lower[f] != c || m->upper[f] != c + X*Y) return false;
                    f++;
                }
            }
    nx = (int32_t)X; ny = (int32_t)Y; nz = (int32_t)Z;
    return true;
}

static void counting_sort(const std::vector<int32_t>& key, int32_t nKeys,
                          std::vector<int32_t>& idx)
{
    // stable counting sort of idx by key[idx[i]]
    std::vector<int64_t> cnt(nKeys + 1, 0);
    for (int32_t i : idx) cnt[key[i] + 1]++;
    for (int32_t k = 0; k < nKeys; k++) cnt[k + 1] += cnt[k];
    std::vector<int32_t> out(idx.size());
    for (int32_t i : idx) out[cnt[key[i]]++] = i;
    idx.swap(out);
}

void build_interfaces(fh_mesh* m, int32_t nInterfaces, const int32_t* start,
                      const int32_t* faceCells, const int32_t* nbrCells)
{
    if (nInterfaces < 0) fail(FH_ERR_INVALID, "nInterfaces must be >= 0");
    if (nInterfaces == 0) return;
    if (!start || !faceCells || !nbrCells) fail(FH_ERR_INVALID, "NULL interface addressing");
    if (start[0] != 0) fail(FH_ERR_INVALID, "interfaceStart[0] must be 0");
    for (int32_t i = 0; i < nInterfaces; i++)
        if (start[i + 1] < start[i])
            fail(FH_ERR_INVALID, "interfaceStart decreases at interface %d", i);
    const int32_t nE = start[nInterfaces];
    const int32_t N = m->nCells;
    for (int32_t e = 0; e < nE; e++)
        if (faceCells[e] < 0 || faceCells[e] >= N || nbrCells[e] < 0 || nbrCells[e] >= N)
            fail(FH_ERR_INVALID, "interface face %d: cells %d / %d out of range", e,
                 faceCells[e], nbrCells[e]);
    m->nInterfaces = nInterfaces;
    m->nIfFaces = nE;
    // per library row, the interface faces in the caller's order (interface
    // by interface, face by face): the order updateMatrixInterfaces adds them
    std::vector<int32_t> cnt(N + 1, 0);
    for (int32_t e = 0; e < nE; e++) cnt[m->iperm[faceCells[e]] + 1]++;
    for (int32_t r = 0; r < N; r++) cnt[r + 1] += cnt[r];
    m->ifStart = cnt;
    m->ifNbr.assign(nE, 0);
    m->ifFace.assign(nE, 0);
    std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
    for (int32_t e = 0; e < nE; e++)
    {
        const int32_t r = m->iperm[faceCells[e]];
        const int32_t k = pos[r]++;
        m->ifNbr[k] = m->iperm[nbrCells[e]];
        m->ifFace[k] = e;
    }
    hipStream_t s = m->setupStream;
    m->d_ifStart.upload(m->ifStart, s);
    m->d_ifNbr.upload(m->ifNbr, s);
    m->d_ifFace.upload(m->ifFace, s);
    FH_HIP(hipStreamSynchronize(s));
}

void build_mesh(fh_mesh* m, int32_t nCells, int32_t nFaces,
                const int32_t* lower, const int32_t* upper, bool allowBox)
{
    if (nCells <= 0) fail(FH_ERR_INVALID, "nCells must be positive (got %d)", nCells);
    if (nFaces < 0) fail(FH_ERR_INVALID, "nFaces must be >= 0 (got %d)", nFaces);
    if (nFaces > 0 && (!lower || !upper))
        fail(FH_ERR_INVALID, "lowerAddr/upperAddr must not be NULL");

    m->nCells = nCells;
    m->nFaces = nFaces;
    m->lower.assign(lower, lower + nFaces);
    m->upper.assign(upper, upper + nFaces);

    // validation: lduAddressing requires l < u and faces sorted by owner
    // (lduAddressing.H:36-89, calcOwnerStart lduAddressing.C:93-127)
    for (int32_t f = 0; f < nFaces; f++)
    {
        const int32_t l = lower[f], u = upper[f];
        if (l < 0 || u >= nCells || l >= u)
            fail(FH_ERR_INVALID,
                 "face %d: lowerAddr %d / upperAddr %d violate 0 <= l < u < nCells",
                 f, l, u);
        if (f > 0 && lower[f] < lower[f - 1])
            fail(FH_ERR_INVALID, "faces are not in upper-triangular (owner) order at face %d", f);
    }

    // levels of the lower-triangular dependency DAG
    m->levelOf.assign(nCells, 0);
    int32_t nLevels = 1;
    for (int32_t f = 0; f < nFaces; f++)
    {
        const int32_t cand = m->levelOf[lower[f]] + 1;
        if (cand > m->levelOf[upper[f]]) m->levelOf[upper[f]] = cand;
    }
    for (int32_t c = 0; c < nCells; c++) nLevels = std::max(nLevels, m->levelOf[c] + 1);

    // partition
    int32_t nx, ny, nz;
    std::vector<int32_t> partOf(nCells, 0);
    int32_t nParts = 1;
    std::vector<int32_t> order(1, 0);
    const char* sweepEnv = std::getenv("FOAMHIP_SWEEP");
    const bool levelSweeps = sweepEnv && std::strcmp(sweepEnv, "levels") == 0;
    // a coupled mesh takes the general rows (the box kernels have no
    // interface terms)
    const bool box = allowBox && detect_box(m, nx, ny, nz);
    if (box && !levelSweeps)
    {
        // natural numbering; the sweeps walk 16 x 16 x bz tiles
        m->boxNx = nx; m->boxNy = ny; m->boxNz = nz;
        m->tiled = true;
    }
    else if (box)
    {
        m->boxNx = nx; m->boxNy = ny; m->boxNz = nz;
        const int T = column_size_for(nx, ny);
        const int32_t nI = (nx + T - 1)/T, nJ = (ny + T - 1)/T;
        nParts = nI*nJ;
        for (int32_t k = 0; k < nz; k++)
            for (int32_t j = 0; j < ny; j++)
                for (int32_t i = 0; i < nx; i++)
                    partOf[i + nx*(j + ny*k)] = (i/T) + nI*(j/T);
        // ticket order: anti-diagonals I+J, a topological order of the
        // column dependency graph (left and lower neighbours come first)
        order.clear();
        for (int32_t d = 0; d <= nI + nJ - 2; d++)
            for (int32_t I = 0; I < nI; I++)
            {
                const int32_t J = d - I;
                if (J >= 0 && J < nJ) order.push_back(I + nI*J);
            }
    }

    // renumber by (part, level, old); tiled boxes keep the natural order
    std::vector<int32_t> idx(nCells);
    for (int32_t c = 0; c < nCells; c++) idx[c] = c;
    if (!m->tiled)
    {
        counting_sort(m->levelOf, nLevels, idx);
        counting_sort(partOf, nParts, idx);
    }
    m->perm = idx;
    m->iperm.assign(nCells, 0);
    for (int32_t r = 0; r < nCells; r++) m->iperm[m->perm[r]] = r;

    // per-cell face lists in reference order
    std::vector<int32_t> loStart(nCells + 1, 0), loFaces(nFaces);
    for (int32_t f = 0; f < nFaces; f++) loStart[upper[f] + 1]++;
    for (int32_t c = 0; c < nCells; c++) loStart[c + 1] += loStart[c];
    {
        std::vector<int32_t> fill(loStart.begin(), loStart.end() - 1);
        for (int32_t f = 0; f < nFaces; f++) loFaces[fill[upper[f]]++] = f;
    }
    std::vector<int32_t> ownStart(nCells + 1, 0);
    for (int32_t f = 0; f < nFaces; f++) ownStart[lower[f] + 1]++;
    for (int32_t c = 0; c < nCells; c++) ownStart[c + 1] += ownStart[c];

    // SELL-64
    const int32_t nSlices = (nCells + 63)/64;
    m->nSlices = nSlices;
    m->sliceStart.assign(nSlices + 1, 0);
    m->rowInfo.assign(nCells, 0);
    for (int32_t r = 0; r < nCells; r++)
    {
        const int32_t c = m->perm[r];
        const int32_t nLo = loStart[c + 1] - loStart[c];
        const int32_t nUp = ownStart[c + 1] - ownStart[c];
        if (nLo + nUp > 65535)
            fail(FH_ERR_UNSUPPORTED, "cell %d has %d neighbours (max 65535)", c, nLo + nUp);
        m->rowInfo[r] = nLo | ((nLo + nUp) << 16);
    }
    int64_t total = 0;
    for (int32_t s = 0; s < nSlices; s++)
    {
        int32_t w = 0;
        for (int32_t r = 64*s; r < std::min(nCells, 64*s + 64); r++)
            w = std::max(w, m->rowInfo[r] >> 16);
        m->sliceStart[s] = (int32_t)total;
        total += 64*(int64_t)w;
        if (total > INT32_MAX) fail(FH_ERR_UNSUPPORTED, "mesh too large for 32-bit entry offsets");
    }
    m->sliceStart[nSlices] = (int32_t)total;
    m->nnzPad = total;
    m->col.assign(total, 0);
    m->face.assign(total, -1);
    for (int32_t r = 0; r < nCells; r++)
    {
        const int32_t c = m->perm[r];
        const int32_t base = m->sliceStart[r >> 6] + (r & 63);
        int32_t k = 0;
        for (int32_t q = loStart[c]; q < loStart[c + 1]; q++, k++)
        {
            const int32_t f = loFaces[q];
            m->col[base + 64*k] = m->iperm[lower[f]];
            m->face[base + 64*k] = (f << 1);          // lower coefficient
        }
        for (int32_t f = ownStart[c]; f < ownStart[c + 1]; f++, k++)
        {
            m->col[base + 64*k] = m->iperm[upper[f]];
            m->face[base + 64*k] = (f << 1) | 1;      // upper coefficient
        }
        // padding: column = own row, never read (k < nEnt test)
        const int32_t w = (m->sliceStart[(r >> 6) + 1] - m->sliceStart[r >> 6])/64;
        for (; k < w; k++) m->col[base + 64*k] = r;
    }

    // schedule
    Schedule& S = m->sched;
    S.nParts = nParts;
    S.nLevels = nLevels;
    std::vector<int32_t> partStart(nParts + 1, 0);
    for (int32_t c = 0; c < nCells; c++) partStart[partOf[c] + 1]++;
    for (int32_t p = 0; p < nParts; p++) partStart[p + 1] += partStart[p];
    S.partLevelOff.assign(nParts + 1, 0);
    S.partMinLevel.assign(nParts, 0);
    S.segStart.clear();
    S.maxSegRows = 0;
    for (int32_t p = 0; p < nParts; p++)
    {
        S.partLevelOff[p] = (int32_t)S.segStart.size();
        const int32_t r0 = partStart[p], r1 = partStart[p + 1];
        if (r0 == r1)
        {
            S.partMinLevel[p] = 0;
            S.segStart.push_back(r0);
            continue;
        }
        const int32_t minL = m->levelOf[m->perm[r0]];
        const int32_t maxL = m->levelOf[m->perm[r1 - 1]];
        S.partMinLevel[p] = minL;
        int32_t r = r0;
        for (int32_t L = minL; L <= maxL; L++)
        {
            S.segStart.push_back(r);
            const int32_t rs = r;
            while (r < r1 && m->levelOf[m->perm[r]] == L) r++;
            S.maxSegRows = std::max(S.maxSegRows, r - rs);
        }
        S.segStart.push_back(r1);
    }
    S.partLevelOff[nParts] = (int32_t)S.segStart.size();

    // part dependencies
    std::vector<std::vector<int32_t>> up(nParts), down(nParts);
    for (int32_t f = 0; f < nFaces; f++)
    {
        const int32_t pl = partOf[lower[f]], pu = partOf[upper[f]];
        if (pl != pu)
        {
            up[pu].push_back(pl);
            down[pl].push_back(pu);
        }
    }
    auto flatten = [&](std::vector<std::vector<int32_t>>& v,
                       std::vector<int32_t>& off, std::vector<int32_t>& lst)
    {
        off.assign(nParts + 1, 0);
        lst.clear();
        for (int32_t p = 0; p < nParts; p++)
        {
            std::sort(v[p].begin(), v[p].end());
            v[p].erase(std::unique(v[p].begin(), v[p].end()), v[p].end());
            off[p] = (int32_t)lst.size();
            lst.insert(lst.end(), v[p].begin(), v[p].end());
        }
        off[nParts] = (int32_t)lst.size();
    };
    flatten(up, S.upOff, S.upList);
    flatten(down, S.downOff, S.downList);
    S.order = order;
    // the ticket order must be topological for both directions, otherwise a
    // waiting part could hold the ticket of the part it waits for
    {
        std::vector<int32_t> pos(nParts);
        for (int32_t t = 0; t < nParts; t++) pos[S.order[t]] = t;
        for (int32_t p = 0; p < nParts; p++)
            for (int32_t q = S.upOff[p]; q < S.upOff[p + 1]; q++)
                if (pos[S.upList[q]] >= pos[p])
                    fail(FH_ERR_INVALID, "internal: sweep order is not topological");
    }

    // chunks of the sync-free sweep: used when the levels hold 16 rows or
    // more on average (narrower dependency graphs, e.g. chains, are walked
    // faster by the one-workgroup level sweep)
    if (!m->tiled && nParts == 1 && (int64_t)nLevels*16 <= (int64_t)nCells)
    {
        std::vector<int32_t> start, chunkOf(nCells);
        for (size_t q = 0; q + 1 < S.segStart.size(); q++)
            for (int32_t r = S.segStart[q]; r < S.segStart[q + 1]; r += 64)
                start.push_back(r);
        start.push_back(nCells);
        const int32_t nCh = (int32_t)start.size() - 1;
        for (int32_t c = 0; c < nCh; c++)
            for (int32_t r = start[c]; r < start[c + 1]; r++) chunkOf[r] = c;
        std::vector<int32_t> fOff(nCh + 1, 0), bOff(nCh + 1, 0), fDeps, bDeps, tmpF, tmpB;
        for (int32_t c = 0; c < nCh; c++)
        {
            tmpF.clear();
            tmpB.clear();
            for (int32_t r = start[c]; r < start[c + 1]; r++)
            {
                const int32_t nLo = m->rowInfo[r] & 0xffff, nEnt = m->rowInfo[r] >> 16;
                const int32_t base = m->sliceStart[r >> 6] + (r & 63);
                for (int32_t k = 0; k < nEnt; k++)
                {
                    const int32_t oc = chunkOf[m->col[base + 64*k]];
                    if (k < nLo)
                    {
                        if (oc >= c) fail(FH_ERR_INVALID, "internal: chunk order is not topological");
                        tmpF.push_back(oc);
                    }
                    else
                    {
                        if (oc <= c) fail(FH_ERR_INVALID, "internal: chunk order is not topological");
                        tmpB.push_back(oc);
                    }
                }
            }
            std::sort(tmpF.begin(), tmpF.end());
            tmpF.erase(std::unique(tmpF.begin(), tmpF.end()), tmpF.end());
            std::sort(tmpB.begin(), tmpB.end());
            tmpB.erase(std::unique(tmpB.begin(), tmpB.end()), tmpB.end());
            fDeps.insert(fDeps.end(), tmpF.begin(), tmpF.end());
            bDeps.insert(bDeps.end(), tmpB.begin(), tmpB.end());
            fOff[c + 1] = (int32_t)fDeps.size();
            bOff[c + 1] = (int32_t)bDeps.size();
        }
        m->chunkSweep = true;
        m->nChunks = nCh;
        m->d_chunkStart.upload(start, m->setupStream);
        m->d_fwdOff.upload(fOff, m->setupStream);
        m->d_fwdDeps.upload(fDeps, m->setupStream);
        m->d_bwdOff.upload(bOff, m->setupStream);
        m->d_bwdDeps.upload(bDeps, m->setupStream);
    }

    // tile shapes and their wavefront order (ascending TI+TJ+TK)
    if (m->tiled)
    {
        // 16 x 16 x {4, 8, 16} (workgroup sweeps), 8 x 8 x 16 (one-wavefront sweeps)
        const int bxs[4] = {16, 16, 16, 8}, bzs[4] = {4, 8, 16, 16};
        for (int t = 0; t < 4; t++)
        {
            fh_mesh::TileShape& T = m->tiles[t];
            T.bx = bxs[t];
            T.bz = bzs[t];
            T.nTI = (nx + T.bx - 1)/T.bx;
            T.nTJ = (ny + T.bx - 1)/T.bx;
            T.nTK = (nz + T.bz - 1)/T.bz;
            std::vector<int32_t> ord;
            ord.reserve((size_t)T.nTI*T.nTJ*T.nTK);
            for (int32_t d = 0; d <= T.nTI + T.nTJ + T.nTK - 3; d++)
                for (int32_t K = 0; K < T.nTK; K++)
                    for (int32_t J = 0; J < T.nTJ; J++)
                    {
                        const int32_t I = d - K - J;
                        if (I >= 0 && I < T.nTI) ord.push_back(I + T.nTI*(J + T.nTJ*K));
                    }
            T.d_order.upload(ord, m->setupStream);
            m->maxTiles = std::max<int32_t>(m->maxTiles, (int32_t)ord.size());
        }
    }

    // upload
    hipStream_t s = m->setupStream;
    m->d_sliceStart.upload(m->sliceStart, s);
    m->d_rowInfo.upload(m->rowInfo, s);
    m->d_col.upload(m->col, s);
    m->d_face.upload(m->face, s);
    m->d_perm.upload(m->perm, s);
    m->d_lower.upload(m->lower, s);
    m->d_upper.upload(m->upper, s);
    S.d_partLevelOff.upload(S.partLevelOff, s);
    S.d_partMinLevel.upload(S.partMinLevel, s);
    S.d_segStart.upload(S.segStart, s);
    S.d_upOff.upload(S.upOff, s);
    S.d_upList.upload(S.upList, s);
    S.d_downOff.upload(S.downOff, s);
    S.d_downList.upload(S.downList, s);
    S.d_order.upload(S.order, s);
    FH_HIP(hipStreamSynchronize(s));
    // host copies of the large SELL arrays are not needed any more
    std::vector<int32_t>().swap(m->col);
    std::vector<int32_t>().swap(m->face);
}

size_t mesh_device_bytes(const fh_mesh* m)
{
    const Schedule& S = m->sched;
    return m->d_sliceStart.bytes() + m->d_rowInfo.bytes() + m->d_col.bytes()
         + m->d_face.bytes() + m->d_perm.bytes() + m->d_lower.bytes()
         + m->d_upper.bytes() + S.d_partLevelOff.bytes() + S.d_partMinLevel.bytes()
         + S.d_segStart.bytes() + S.d_upOff.bytes() + S.d_upList.bytes()
         + S.d_downOff.bytes() + S.d_downList.bytes() + S.d_order.bytes()
         + m->d_chunkStart.bytes() + m->d_fwdOff.bytes() + m->d_fwdDeps.bytes()
         + m->d_bwdOff.bytes() + m->d_bwdDeps.bytes() + m->d_ifStart.bytes()
         + m->d_ifNbr.bytes() + m->d_ifFace.bytes();
}

} // namespace fh
