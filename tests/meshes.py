"""Small test meshes: hex boxes in blockMesh order and random LDU graphs in
upper-triangular face order (owner ascending; neighbours within an owner
either ascending, as polyMesh has them, or shuffled, as GAMG coarse levels
have them)."""
import numpy as np


def box(nx, ny, nz):
    lower, upper = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                if i < nx - 1:
                    lower.append(c); upper.append(c + 1)
                if j < ny - 1:
                    lower.append(c); upper.append(c + nx)
                if k < nz - 1:
                    lower.append(c); upper.append(c + nx * ny)
    return nx * ny * nz, np.array(lower, np.int32), np.array(upper, np.int32)


def random_graph(n, avg_deg, rng, shuffle_within_owner=False, connect_chain=True):
    """Random symmetric sparsity, faces sorted by owner."""
    pairs = set()
    if connect_chain:
        for c in range(n - 1):
            pairs.add((c, c + 1))
    m = int(n * avg_deg / 2)
    while len(pairs) < m + (n - 1 if connect_chain else 0):
        a, b = rng.integers(0, n, 2)
        if a == b:
            continue
        pairs.add((min(a, b), max(a, b)))
    pairs = sorted(pairs)
    lower = np.array([p[0] for p in pairs], np.int32)
    upper = np.array([p[1] for p in pairs], np.int32)
    if shuffle_within_owner:
        # keep owner order, permute faces inside each owner block
        idx = np.arange(len(lower))
        out = []
        start = 0
        while start < len(lower):
            end = start
            while end < len(lower) and lower[end] == lower[start]:
                end += 1
            blk = idx[start:end].copy()
            rng.shuffle(blk)
            out.extend(blk)
            start = end
        out = np.array(out)
        lower, upper = lower[out], upper[out]
    return n, lower, upper


def hub_graph(n, avg_deg, n_hubs, hub_deg, rng):
    """A random graph (no chain) plus hub cells, each joined to hub_deg
    random cells below it and hub_deg above it: rows with more lower and
    more upper neighbours than a sweep keeps in registers."""
    pairs = set()
    m = int(n * avg_deg / 2)
    while len(pairs) < m:
        a, b = rng.integers(0, n, 2)
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    for h in rng.choice(np.arange(n // 4, 3 * n // 4), n_hubs, replace=False):
        for c in rng.choice(h, hub_deg, replace=False):
            pairs.add((int(c), int(h)))
        for c in h + 1 + rng.choice(n - h - 1, hub_deg, replace=False):
            pairs.add((int(h), int(c)))
    pairs = sorted(pairs)
    return (n, np.array([p[0] for p in pairs], np.int32),
            np.array([p[1] for p in pairs], np.int32))


def spd_coeffs(n, lower, upper, rng, negative=False):
    """Diagonally dominant symmetric coefficients (M-matrix pattern)."""
    up = -rng.uniform(0.1, 1.0, len(lower))
    diag = np.zeros(n)
    np.subtract.at(diag, lower, up)
    np.subtract.at(diag, upper, up)
    diag += rng.uniform(0.01, 0.5, n)
    if negative:
        return -diag, -up
    return diag, up


def asym_coeffs(n, lower, upper, rng):
    up = -rng.uniform(0.1, 1.0, len(lower))
    lo = -rng.uniform(0.1, 1.0, len(lower))
    diag = np.zeros(n)
    np.subtract.at(diag, lower, lo)
    np.subtract.at(diag, upper, up)
    diag += rng.uniform(0.05, 0.5, n)
    return diag, up, lo


def dense(n, lower, upper, diag, up, lo=None):
    lo = up if lo is None else lo
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = diag
    A[lower, upper] = up
    A[upper, lower] = lo
    return A


def cyclic_box(nx, ny, nz):
    """A hex box whose x-min and x-max patches form one cyclic pair: the
    LDU addressing of the box, and the two coupled interfaces (x-min
    patch first) with their faces in (k, j) order, face e of one patch
    matching face e of the other."""
    n, lower, upper = box(nx, ny, nz)
    lo = np.array([nx * (j + ny * k) for k in range(nz) for j in range(ny)], np.int32)
    hi = lo + (nx - 1)
    nf = len(lo)
    return n, lower, upper, {"start": np.array([0, nf, 2 * nf], np.int32),
                             "faceCells": np.concatenate([lo, hi]),
                             "nbrCells": np.concatenate([hi, lo])}


def dense_interfaces(n, ifc, coeffs):
    """The interfaces' part of the dense matrix: result[faceCells] -=
    coeffs*psi[nbrCells] is A[faceCells, nbrCells] -= coeffs."""
    C = np.zeros((n, n))
    np.add.at(C, (ifc["faceCells"], ifc["nbrCells"]), -np.asarray(coeffs))
    return C


def cyclic_system(kind, seed=3, shape=(6, 5, 4), with_dense=True):
    """Random coefficients on a box with an x cyclic pair: symmetric
    ('sym', the same coefficient on both faces of a pair) or asymmetric
    (interfaceIntCoeffs of a face = bouCoeffs of its partner face, so that
    Tmul is the transpose); the couplings -bou have the sign of the internal
    off-diagonals and their magnitude is added to the diagonal.  Returns the
    addressing, coefficients, interfaces and the dense matrix and its
    Tmul counterpart."""
    rng = np.random.default_rng(seed)
    n, lower, upper, ifc = cyclic_box(*shape)
    nf = ifc["start"][1]
    if kind == "sym":
        diag, up = spd_coeffs(n, lower, upper, rng)
        lo = None
        # the same coefficient on both sides of a cyclic face pair
        b0 = rng.uniform(0.1, 1.0, nf)
        bou = np.concatenate([b0, b0])
        itc = bou
    else:
        diag, up, lo = asym_coeffs(n, lower, upper, rng)
        bou = rng.uniform(0.1, 1.0, 2 * nf)
        # the transpose of the coupling: face e of one patch carries as
        # interfaceIntCoeffs the boundary coefficient of face e of the other
        itc = np.concatenate([bou[nf:], bou[:nf]])
    # the coupled faces' couplings -bou have the sign of the internal
    # off-diagonal coefficients (negative here); their magnitude goes on the
    # diagonal, as addBoundaryDiag does with the coupled patch's
    # internalCoeffs, keeping the matrix diagonally dominant
    np.add.at(diag, ifc["faceCells"], bou + 0.05)
    ifc = dict(ifc, bouCoeffs=bou, intCoeffs=itc)
    if not with_dense:
        return n, lower, upper, diag, up, lo, ifc, None, None
    A = dense(n, lower, upper, diag, up, lo) + dense_interfaces(n, ifc, bou)
    AT = dense(n, lower, upper, diag, up, lo).T + dense_interfaces(n, ifc, itc)
    return n, lower, upper, diag, up, lo, ifc, A, AT


def renumber_cells(N, lower, upper, faceStart, facePoints, q):
    """The same mesh with cell c renumbered q[c]: internal faces re-oriented
    so that the lower new index owns them (a flipped face has its points
    reversed, so its normal still points from owner to neighbour) and
    re-sorted into upper-triangular order.  Returns the new lower, upper,
    faceStart, facePoints and, per new internal face, its old index."""
    q = np.asarray(q)
    l, u = q[lower], q[upper]
    flip = l > u
    nl, nu = np.where(flip, u, l), np.where(flip, l, u)
    order = np.lexsort((nu, nl))
    F = len(lower)
    pts = [facePoints[faceStart[f]:faceStart[f + 1]] for f in range(len(faceStart) - 1)]
    new_pts = []
    for f in order:
        p = pts[f]
        new_pts.append(p[::-1] if flip[f] else p)
    new_pts += pts[F:]
    fs = np.zeros(len(new_pts) + 1, np.int32)
    fs[1:] = np.cumsum([len(p) for p in new_pts])
    return (nl[order].astype(np.int32), nu[order].astype(np.int32), fs,
            np.concatenate(new_pts).astype(np.int32), order)
