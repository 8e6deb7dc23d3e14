"""Independent restatements used to pin the oracle (pure Python / numpy)."""
import numpy as np


def morton_key(loc):
    """87-bit Morton key of 29-bit codes: triple (x + 2y + 4z) per level,
    level 28 most significant (child index of src/octree.c:101-107)."""
    x, y, z = (int(v) for v in loc)
    k = 0
    for lvl in range(28, -1, -1):
        k = (k << 3) | (((x >> lvl) & 1) | (((y >> lvl) & 1) << 1) | (((z >> lvl) & 1) << 2))
    return k


def msb(v):
    return int(v).bit_length() - 1


def brute_leaf_levels(loc):
    """level of the PR-octree leaf of each cell: the level at which it first
    parts from every other cell, min over r != q of msb(OR_a(q_a ^ r_a))."""
    loc = np.asarray(loc, np.uint64)
    n = len(loc)
    if n == 1:
        return np.array([29])
    x = loc[:, 0] ^ loc[:, 0][:, None]
    y = loc[:, 1] ^ loc[:, 1][:, None]
    z = loc[:, 2] ^ loc[:, 2][:, None]
    orr = (x | y | z).astype(np.float64)
    np.fill_diagonal(orr, np.inf)
    m = orr.min(axis=1)
    return np.floor(np.log2(m)).astype(np.int64)


def leaf_octants(loc, levels):
    loc = np.asarray(loc, np.int64)
    lv = np.asarray(levels, np.int64)
    mask = ~((np.int64(1) << lv) - 1)
    lo = loc & mask[:, None]
    hi = lo + (np.int64(1) << lv)[:, None]
    return lo, hi


def textbook_w(r, h):
    """M4 cubic spline with compact support h (Monaghan 1992; Springel 2010 eq. 4)"""
    q = r / h
    s = 8.0 / (np.pi * h ** 3)
    return np.where(q <= 0.5, s * (1 - 6 * q ** 2 + 6 * q ** 3), np.where(q <= 1.0, s * 2 * (1 - q) ** 3, 0.0))


def lens_volume(r1, r2, d):
    """volume of the intersection of two spheres (textbook closed form)"""
    if d >= r1 + r2:
        return 0.0
    return np.pi * (r1 + r2 - d) ** 2 * (d * d + 2 * d * (r1 + r2) - 3 * (r1 - r2) ** 2) / (12 * d)
