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
