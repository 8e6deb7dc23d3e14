"""Seeded synthetic colonies and grids for the cell-mechanics path.

Counter-based SplitMix64 streams (never libc rand), generated on the host in
FP64 so the CPU checker and the GPU path read identical bytes.  Cell
parameters follow the reference's own laws where it has one
(young = 2100 + 100*U(-1,1), src/cells.c:187 of mcytu/timothy).

Workloads (seed = 1000 + index):
  0  uniform ball, 20 000 cells, r=1, h=2, volume fraction 0.6, 64^3 grid
  1  Gaussian spheroid, 2 000 000 cells, sigma=0.35*W (z*0.7), W=146.6, 256^3 grid
  2  the same cells as 1 (mixed-precision form)
  3  256 colonies, N_i log-uniform in [1e3, 2e5], ball or Gaussian by coin flip
  4  Gaussian spheroid, 50 000 000 cells, 5% radius 1.5, 512^3 FP32 field
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

TYPE0 = dict(size=1.0, h=2.0, h2=4.0, h3=8.0, h4=16.0)
TYPE1 = dict(size=1.5, h=3.0, h2=9.0, h3=27.0, h4=81.0)
CELLTYPES = [TYPE0, TYPE1]


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """SplitMix64 output number idx of the stream started at `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """A SplitMix64 stream read in consecutive blocks."""

    def __init__(self, seed: int, sub: int = 0):
        self.seed = int(splitmix64(seed, np.array([sub], np.uint64))[0])
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        out = splitmix64(self.seed, np.arange(self.pos, self.pos + n, dtype=np.uint64))
        self.pos += n
        return out

    def uniform(self, n: int) -> np.ndarray:
        """doubles in [0, 1) with 53 random bits"""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller normals"""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)])
        return z[:n]


def uniform_ball(n: int, radius: float, st: Stream) -> np.ndarray:
    pts = []
    have = 0
    while have < n:
        m = int((n - have) * 2.0) + 64
        c = (st.uniform(3 * m).reshape(m, 3) * 2.0 - 1.0) * radius
        keep = c[(c * c).sum(1) <= radius * radius]
        pts.append(keep)
        have += len(keep)
    return np.concatenate(pts)[:n]


def gaussian_spheroid(n: int, half_width: float, st: Stream, sigma_frac=0.35, z_ratio=0.7) -> np.ndarray:
    s = sigma_frac * half_width
    sig = np.array([s, s, s * z_ratio])
    pts = []
    have = 0
    while have < n:
        m = int((n - have) * 1.1) + 64
        c = st.normal(3 * m).reshape(m, 3) * sig
        keep = c[np.all(np.abs(c) <= half_width, axis=1)]
        pts.append(keep)
        have += len(keep)
    return np.concatenate(pts)[:n]


def _cells(pos: np.ndarray, ctype: np.ndarray, st: Stream, types=CELLTYPES) -> dict:
    n = len(pos)
    sizes = np.array([t["size"] for t in types])
    young = (2100.0 + 100.0 * (st.uniform(n) * 2.0 - 1.0)).astype(np.float32)
    return dict(
        n=n,
        x=np.ascontiguousarray(pos[:, 0], np.float64),
        y=np.ascontiguousarray(pos[:, 1], np.float64),
        z=np.ascontiguousarray(pos[:, 2], np.float64),
        size=sizes[ctype].astype(np.float64),
        young=young,
        ctype=ctype.astype(np.int32),
        phase=np.ones(n, np.int32),
        types=types,
    )


def ball_colony(n: int, seed: int, volume_fraction=0.6, radius=1.0) -> dict:
    st = Stream(seed, 1)
    R = radius * (n / volume_fraction) ** (1.0 / 3.0)
    pos = uniform_ball(n, R, st)
    d = _cells(pos, np.zeros(n, np.int32), Stream(seed, 2))
    d["shape"] = ("ball", R)
    return d


def gaussian_colony(n: int, seed: int, half_width: float, frac_type1=0.0) -> dict:
    st = Stream(seed, 1)
    pos = gaussian_spheroid(n, half_width, st)
    if frac_type1 > 0:
        ctype = (Stream(seed, 3).uniform(n) < frac_type1).astype(np.int32)
    else:
        ctype = np.zeros(n, np.int32)
    d = _cells(pos, ctype, Stream(seed, 2))
    d["shape"] = ("gaussian", half_width)
    return d


def bbox_grid(col: dict, G: int) -> dict:
    """G^3 grid spanning the cells: lower = min(pos), float resolution rounded
    up from max_extent/(G-2) so every cell's eight CIC corners are in range."""
    lower = np.array([col["x"].min(), col["y"].min(), col["z"].min()])
    upper = np.array([col["x"].max(), col["y"].max(), col["z"].max()])
    ext = float((upper - lower).max())
    want = ext / (G - 2) if ext > 0 else 1.0
    res = np.float32(want)
    if float(res) < want:
        res = np.nextafter(res, np.float32(np.inf))
    return dict(lower=lower, resolution=np.float32(res), loweridx=np.zeros(3, np.int64),
                upperidx=np.full(3, G - 1, np.int64), localsize=np.full(3, G, np.int64), G=G)


def config(index: int, scale: float = 1.0) -> dict:
    """Workload `index` (0..4).  `scale` < 1 shrinks cell counts keeping density
    (for quick tests); scale=1 is the named configuration."""
    seed = 1000 + index
    if index == 0:
        n = max(2, int(20000 * scale))
        col = ball_colony(n, seed)
        col["grid"] = bbox_grid(col, 64)
    elif index in (1, 2):
        n = max(2, int(2_000_000 * scale))
        W = 146.6 * scale ** (1.0 / 3.0)
        col = gaussian_colony(n, 1001, W)
        col["grid"] = bbox_grid(col, 256)
    elif index == 4:
        n = max(2, int(50_000_000 * scale))
        W = 146.6 * (25.0 * scale) ** (1.0 / 3.0)
        col = gaussian_colony(n, seed, W, frac_type1=0.05)
        col["grid"] = bbox_grid(col, 512)
    else:
        raise ValueError("use colonies_config3 for the batched workload")
    col["config"] = index
    col["seed"] = seed if index != 2 else 1001
    return col


def colonies_config3(ncol: int = 256, scale: float = 1.0, seed: int = 1003) -> list:
    """Batched workload: ncol colonies, N_i = floor(exp(U(ln 1e3, ln 2e5))) (times
    `scale`), uniform ball (phi 0.6) or Gaussian spheroid by seeded coin flip."""
    st = Stream(seed, 7)
    u = st.uniform(ncol)
    flip = st.uniform(ncol)
    cols = []
    for i in range(ncol):
        n = int(math.floor(math.exp(math.log(1e3) + u[i] * (math.log(2e5) - math.log(1e3)))))
        n = max(2, int(n * scale))
        cseed = seed * 1000003 + i
        if flip[i] < 0.5:
            c = ball_colony(n, cseed)
        else:
            c = gaussian_colony(n, cseed, 146.6 * (n / 2e6) ** (1.0 / 3.0))
        cols.append(c)
    return cols


def smooth_field(grid: dict, seed: int, dtype=np.float32) -> np.ndarray:
    """Sum of 8 seeded Gaussians A_k exp(-|x-mu_k|^2 / (2 s_k^2)) sampled on the
    grid nodes (k fastest), A_k~U(0.5,1.5), mu_k uniform in the box,
    s_k~U(0.05,0.2)*extent."""
    st = Stream(seed, 11)
    G = int(grid["G"])
    res = float(grid["resolution"])
    lower = grid["lower"]
    extent = res * (G - 1)
    A = 0.5 + st.uniform(8)
    mu = lower + st.uniform(24).reshape(8, 3) * extent
    s = (0.05 + 0.15 * st.uniform(8)) * extent
    ax = [lower[a] + res * np.arange(G) for a in range(3)]
    out = np.zeros((G, G, G), np.float64)
    for k in range(8):
        gx = np.exp(-((ax[0] - mu[k, 0]) ** 2) / (2 * s[k] ** 2))
        gy = np.exp(-((ax[1] - mu[k, 1]) ** 2) / (2 * s[k] ** 2))
        gz = np.exp(-((ax[2] - mu[k, 2]) ** 2) / (2 * s[k] ** 2))
        out += A[k] * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    return np.ascontiguousarray(out.reshape(-1), dtype)
