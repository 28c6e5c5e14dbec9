"""Seeded synthetic catalogues for tests and benchmarks.

All generators use numpy's PCG64 (`np.random.default_rng(seed)`) and return
float64 arrays, points as (n, 3) in a periodic cube [0, box)^3.
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, box: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, box, size=(n, 3))


def neyman_scott(n_poisson: int, n_parents: int, members: int, sigma: float, box: float,
                 seed: int) -> np.ndarray:
    """n_poisson uniform points plus n_parents clusters of `members` points drawn
    from an isotropic Gaussian of width sigma around a uniform parent, wrapped
    into the box.  Total n_poisson + n_parents * members points, shuffled."""
    rng = np.random.default_rng(seed)
    pois = rng.uniform(0.0, box, size=(n_poisson, 3))
    parents = rng.uniform(0.0, box, size=(n_parents, 3))
    off = rng.normal(0.0, sigma, size=(n_parents, members, 3))
    clus = np.mod(parents[:, None, :] + off, box).reshape(-1, 3)
    pts = np.concatenate([pois, clus], axis=0)
    rng.shuffle(pts, axis=0)
    return pts


def clustered(n: int, box: float, seed: int, members: int = 20, sigma: float = 2.0,
              poisson_frac: float = 0.5) -> np.ndarray:
    """The clustered recipe used throughout: poisson_frac of the points uniform,
    the rest in Neyman-Scott clusters of `members` points with Gaussian sigma."""
    n_clu = int(round(n * (1.0 - poisson_frac)))
    n_par = n_clu // members
    n_poi = n - n_par * members
    return neyman_scott(n_poi, n_par, members, sigma, box, seed)


def lognormal_weights(n: int, sigma: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, size=n))


def grf_templates(n_fields: int, grid: int, box: float, corr_len: float, seed: int) -> np.ndarray:
    """n_fields unit-variance Gaussian random fields on a periodic grid^3 mesh
    with power spectrum P(k) ~ exp(-k^2 l^2), returned as (n_fields, grid, grid, grid) float64."""
    rng = np.random.default_rng(seed)
    k1 = 2.0 * np.pi * np.fft.fftfreq(grid, d=box / grid)
    k1r = 2.0 * np.pi * np.fft.rfftfreq(grid, d=box / grid)
    kx, ky, kz = np.meshgrid(k1, k1, k1r, indexing="ij")
    amp = np.sqrt(np.exp(-(kx * kx + ky * ky + kz * kz) * corr_len * corr_len))
    out = np.empty((n_fields, grid, grid, grid))
    for f in range(n_fields):
        white = rng.normal(size=(grid, grid, grid))
        field = np.fft.irfftn(np.fft.rfftn(white) * amp, s=(grid, grid, grid), axes=(0, 1, 2))
        field -= field.mean()
        field /= field.std()
        out[f] = field
    return out


def sample_trilinear(fields: np.ndarray, xyz: np.ndarray, box: float) -> np.ndarray:
    """Periodic trilinear interpolation of (F, g, g, g) fields at points (n, 3) -> (n, F)."""
    g = fields.shape[1]
    u = np.mod(xyz / box * g, g)
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    i1 = (i0 + 1) % g
    i0 %= g
    out = np.zeros((xyz.shape[0], fields.shape[0]))
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1.0 - f[:, 0]
        ix = i1[:, 0] if dx else i0[:, 0]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1.0 - f[:, 1]
            iy = i1[:, 1] if dy else i0[:, 1]
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1.0 - f[:, 2]
                iz = i1[:, 2] if dz else i0[:, 2]
                out += (wx * wy * wz)[:, None] * fields[:, ix, iy, iz].T
    return out


def lattice(n_side: int, spacing: float = 1.0, origin: float = 0.0) -> np.ndarray:
    """Points of a cubic lattice: every pair distance is sqrt(integer) * spacing,
    so bin edges at integer radii are hit exactly."""
    a = origin + spacing * np.arange(n_side, dtype=np.float64)
    x, y, z = np.meshgrid(a, a, a, indexing="ij")
    return np.column_stack([x.ravel(), y.ravel(), z.ravel()])


def regions(n_regions: int, seed: int, rmax: float, median: float = 4e4, sigma: float = 1.0,
            n_min: float = 5e3, n_max: float = 4e5, density: float = 5e-4, ran_factor: int = 2):
    """Domain-decomposed sub-volumes of uneven size.  Region r has N_r data
    points (lognormal around `median`, clipped to [n_min, n_max]) in a cube of
    side (N_r/density)^(1/3), clustered by the recipe of clustered(), plus
    ran_factor*N_r uniform randoms, and an rmax-wide shell around the cube at
    the same densities whose points are flagged as buffer (never an apex).
    Returns xyz (n,3), buf (n,) int32, off (n_regions+1,) int64 and
    ndata (n_regions,) int64; region r is xyz[off[r]:off[r+1]], data first."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(median), sigma, n_regions)), n_min, n_max).astype(np.int64)
    xs, bufs, off, ndata = [], [], [0], []
    for r, nr in enumerate(sizes):
        side = (nr / density) ** (1.0 / 3.0)
        big = side + 2.0 * rmax
        f = (big / side) ** 3
        nd = int(round(nr * f))
        d = clustered(nd, big, seed * 1000 + 2 * r) - rmax
        ran = uniform(int(round(ran_factor * nr * f)), big, seed * 1000 + 2 * r + 1) - rmax
        pts = np.concatenate([d, ran])
        inside = np.all((pts >= 0.0) & (pts < side), axis=1)
        # place the regions side by side (they never interact)
        pts = pts + np.array([r * (big + 10.0), 0.0, 0.0])
        xs.append(pts)
        bufs.append((~inside).astype(np.int32))
        ndata.append(len(d))
        off.append(off[-1] + len(pts))
    return (np.concatenate(xs), np.concatenate(bufs), np.array(off, dtype=np.int64),
            np.array(ndata, dtype=np.int64))
