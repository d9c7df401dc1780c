"""Reader of OpenFOAM's `constant/polyMesh` (ASCII, optionally gzipped):
points, faces, owner, neighbour and boundary, as written by the reference's
`polyMesh::write` (`src/OpenFOAM/meshes/polyMesh/polyMesh.C`,
`polyMeshIO.C`) through `IOobject` headers (`FoamFile { ... }`) and the list
format of `ListIO.C` / `UListIO.C`: an optional size, then `(` items `)`,
or `N{value}` for a uniform list.  Faces are `n(p0 p1 ...)`
(`faceList`; `faceCompactList` as two label lists is read as well).

The mesh is polyMesh's: faces 0 .. nInternalFaces-1 are internal (owner <
neighbour, owners ascending), the rest belong to the boundary patches in
`boundary` order (`startFace`, `nFaces`); a neighbour list that carries -1
for the boundary faces (older files) is cut to the internal ones.  `lduAddressing` is
(owner, neighbour) of the internal faces (`fvMeshLduAddressing.H:74-92`).

Binary-format files are refused (`format binary;` in the header).
"""
import gzip
import os
import re

import numpy as np

_COMMENT = re.compile(r"//[^\n]*|/\*.*?\*/", re.S)


def _read_text(path):
    if os.path.exists(path):
        with open(path, "rb") as f:
            data = f.read()
    elif os.path.exists(path + ".gz"):
        with gzip.open(path + ".gz", "rb") as f:
            data = f.read()
    else:
        raise FileNotFoundError(path)
    return data.decode("latin-1")


def _split_header(text):
    """(header dict entries, body) of a file with a FoamFile header."""
    text = _COMMENT.sub(" ", text)
    m = re.search(r"FoamFile\s*\{(.*?)\}", text, re.S)
    header = {}
    body = text
    if m:
        for key, val in re.findall(r"(\w+)\s+([^;]*);", m.group(1)):
            header[key] = val.strip().strip('"')
        body = text[m.end():]
    if header.get("format", "ascii") != "ascii":
        raise ValueError("binary polyMesh files are not read: write the mesh with "
                         "'writeFormat ascii'")
    return header, body


def _label_list(body):
    body = body.strip()
    m = re.match(r"(\d+)\s*\{\s*(-?\d+)\s*\}", body)
    if m:
        return np.full(int(m.group(1)), int(m.group(2)), dtype=np.int32)
    m = re.match(r"(\d+)\s*\(", body)
    if m:
        n = int(m.group(1))
        inner = body[m.end():body.rindex(")")]
    else:
        n = None
        inner = body[body.index("(") + 1:body.rindex(")")]
    out = np.array(inner.split(), dtype=np.int64).astype(np.int32)
    if n is not None and out.size != n:
        raise ValueError(f"list of {n} labels holds {out.size}")
    return out


def _vector_list(body):
    body = body.strip()
    m = re.match(r"(\d+)\s*\(", body)
    n = int(m.group(1)) if m else None
    start = m.end() if m else body.index("(") + 1
    inner = body[start:body.rindex(")")]
    vals = np.array(inner.replace("(", " ").replace(")", " ").split(), dtype=np.float64)
    out = vals.reshape(-1, 3)
    if n is not None and out.shape[0] != n:
        raise ValueError(f"list of {n} vectors holds {out.shape[0]}")
    return out


def _face_list(body, header):
    """(faceStart[nFaces+1], facePoints) of a faceList or faceCompactList."""
    if header.get("class") == "faceCompactList":
        parts = body.strip()
        first = parts.index("(")
        depth, i = 0, first
        while True:
            if parts[i] == "(":
                depth += 1
            elif parts[i] == ")":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        offsets = _label_list(parts[:i + 1])
        points = _label_list(parts[i + 1:])
        return offsets.astype(np.int32), points.astype(np.int32)
    body = body.strip()
    m = re.match(r"(\d+)\s*\(", body)
    n = int(m.group(1))
    inner = body[m.end():body.rindex(")")]
    sizes = []
    pts = []
    for sz, lst in re.findall(r"(\d+)\s*\(([^)]*)\)", inner):
        items = lst.split()
        if len(items) != int(sz):
            raise ValueError("face size does not match its point list")
        sizes.append(int(sz))
        pts.append(items)
    if len(sizes) != n:
        raise ValueError(f"list of {n} faces holds {len(sizes)}")
    start = np.zeros(n + 1, dtype=np.int32)
    start[1:] = np.cumsum(sizes)
    flat = np.array([p for f in pts for p in f], dtype=np.int64).astype(np.int32)
    return start, flat


def _boundary(body):
    """[(name, type, startFace, nFaces)] in file order."""
    body = body.strip()
    inner = body[body.index("(") + 1:body.rindex(")")]
    patches = []
    for name, entries in re.findall(r"(\w+)\s*\{([^}]*)\}", inner):
        d = dict(re.findall(r"(\w+)\s+([^;]*);", entries))
        patches.append((name, d.get("type", "patch").strip(), int(d["startFace"]),
                        int(d["nFaces"])))
    return patches


class PolyMesh:
    """A polyMesh read from `<case>/constant/polyMesh` (or the directory
    itself): points [nPoints, 3], faceStart [nFaces+1] / facePoints, owner
    [nFaces], neighbour [nInternalFaces], patches [(name, type, startFace,
    nFaces)]; nCells = max(owner, neighbour) + 1."""

    def __init__(self, path):
        d = path
        if os.path.isdir(os.path.join(path, "constant", "polyMesh")):
            d = os.path.join(path, "constant", "polyMesh")
        h, b = _split_header(_read_text(os.path.join(d, "points")))
        self.points = _vector_list(b)
        h, b = _split_header(_read_text(os.path.join(d, "faces")))
        self.faceStart, self.facePoints = _face_list(b, h)
        h, b = _split_header(_read_text(os.path.join(d, "owner")))
        self.owner = _label_list(b)
        h, b = _split_header(_read_text(os.path.join(d, "neighbour")))
        self.neighbour = _label_list(b)
        h, b = _split_header(_read_text(os.path.join(d, "boundary")))
        self.patches = _boundary(b)
        self._check()

    @classmethod
    def from_arrays(cls, points, faceStart, facePoints, owner, neighbour, patches):
        m = cls.__new__(cls)
        m.points = np.asarray(points, dtype=np.float64)
        m.faceStart = np.asarray(faceStart, dtype=np.int32)
        m.facePoints = np.asarray(facePoints, dtype=np.int32)
        m.owner = np.asarray(owner, dtype=np.int32)
        m.neighbour = np.asarray(neighbour, dtype=np.int32)
        m.patches = [(str(n), str(t), int(s), int(c)) for n, t, s, c in patches]
        m._check()
        return m

    def _check(self):
        nF = self.faceStart.size - 1
        # older files list a neighbour for every face, -1 on the boundary
        neg = np.nonzero(self.neighbour < 0)[0]
        if neg.size:
            if neg[0] + neg.size != self.neighbour.size:
                raise ValueError("neighbour: -1 entries before internal faces")
            self.neighbour = self.neighbour[:neg[0]]
        if self.owner.size != nF:
            raise ValueError(f"owner has {self.owner.size} entries for {nF} faces")
        self.nFaces = nF
        self.nInternalFaces = self.neighbour.size
        self.nCells = int(max(self.owner.max(), self.neighbour.max() if self.neighbour.size
                              else -1)) + 1
        if self.patches:
            end = self.patches[-1][2] + self.patches[-1][3]
            if self.patches[0][2] != self.nInternalFaces or end != nF:
                raise ValueError("the boundary patches do not cover the boundary faces")
        if self.facePoints.size and (self.facePoints.min() < 0
                                     or self.facePoints.max() >= self.points.shape[0]):
            raise ValueError("a face refers to a point that does not exist")

    @property
    def lower(self):
        """lduAddressing lowerAddr: the owner of each internal face."""
        return self.owner[:self.nInternalFaces]

    @property
    def upper(self):
        """lduAddressing upperAddr: the neighbour of each internal face."""
        return self.neighbour

    @property
    def bFaceCells(self):
        """The cell of each boundary face, in patch order."""
        return self.owner[self.nInternalFaces:]

    def patch_of_boundary_faces(self):
        """Index into `patches` of each boundary face."""
        out = np.empty(self.nFaces - self.nInternalFaces, dtype=np.int32)
        for i, (_, _, start, n) in enumerate(self.patches):
            out[start - self.nInternalFaces:start - self.nInternalFaces + n] = i
        return out
