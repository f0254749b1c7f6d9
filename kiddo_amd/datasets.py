"""Seeded synthetic point sets for the benchmark workloads and parity tests.

Every generator is a pure function of (sizes, seed) on numpy's PCG64 stream, so the
GPU run, the CPU baseline and the tests all see identical inputs.
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, dims: int = 3, seed: int = 0, dtype=np.float64) -> np.ndarray:
    """n points uniform in [0, 1)^dims."""
    rng = np.random.default_rng(seed)
    return rng.random((n, dims)).astype(dtype)


def uniform_with_duplicates(n: int, dup_frac: float = 0.05, dims: int = 3, seed: int = 0,
                            dtype=np.float64) -> np.ndarray:
    """n uniform points followed by round(dup_frac * n) exact copies of uniformly chosen
    earlier points (copies carry item ids >= n), forcing distance ties."""
    rng = np.random.default_rng(seed)
    base = rng.random((n, dims)).astype(dtype)
    ndup = int(round(dup_frac * n))
    src = rng.integers(0, n, size=ndup)
    return np.concatenate([base, base[src]], axis=0)


def gaussian_mixture(n: int, components: int = 64, sigma: float = 0.01, dims: int = 3,
                     seed: int = 0, dtype=np.float32, centres=None) -> np.ndarray:
    """Isotropic Gaussian mixture in the unit cube: centres uniform, equal weights, points
    clipped to [0, 1)."""
    rng = np.random.default_rng(seed)
    if centres is None:
        centres = np.random.default_rng(seed ^ 0x5EED).random((components, dims))
    comp = rng.integers(0, len(centres), size=n)
    pts = centres[comp] + rng.normal(0.0, sigma, size=(n, dims))
    np.clip(pts, 0.0, np.nextafter(1.0, 0.0), out=pts)
    return pts.astype(dtype)


def noisy_sphere(n: int, seed: int = 0, jitter: float = 0.01, dtype=np.float32) -> np.ndarray:
    """Unit-sphere directions scaled by (1 + N(0, jitter))."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v *= 1.0 + rng.normal(0.0, jitter, size=(n, 1))
    return v.astype(dtype)


def latlon_clusters(n: int, clusters: int = 200, cluster_frac: float = 0.7,
                    sigma_deg: float = 0.5, seed: int = 0, cluster_seed: int = 77) -> np.ndarray:
    """Unit vectors: cluster_frac of the mass from Gaussian lat/lon clusters, the rest
    uniform on the sphere."""
    rng = np.random.default_rng(seed)
    crng = np.random.default_rng(cluster_seed)
    clat = np.degrees(np.arcsin(crng.uniform(-1, 1, clusters)))
    clon = crng.uniform(-180, 180, clusters)
    nc = int(round(cluster_frac * n))
    which = rng.integers(0, clusters, size=nc)
    lat = np.clip(clat[which] + rng.normal(0, sigma_deg, nc), -90, 90)
    lon = clon[which] + rng.normal(0, sigma_deg, nc)
    lat_u = np.degrees(np.arcsin(rng.uniform(-1, 1, n - nc)))
    lon_u = rng.uniform(-180, 180, n - nc)
    lat = np.radians(np.concatenate([lat, lat_u]))
    lon = np.radians(np.concatenate([lon, lon_u]))
    xyz = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)
    return xyz.astype(np.float64)


def palette_u16(n: int = 65536, dims: int = 4, seed: int = 0) -> np.ndarray:
    """n colours with u16 channels in [0, 65534] (65535 is Kiddo's no-pivot sentinel)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 65535, size=(n, dims), dtype=np.uint16)


def weight_rank_items(n: int, seed: int = 0) -> np.ndarray:
    """Item ids that encode a seeded per-item f32 weight: item = rank of the weight, so the
    smallest item (Kiddo's best_n_within "best") is the entry with the smallest weight."""
    w = np.random.default_rng(seed).random(n, dtype=np.float32)
    order = np.argsort(w, kind="stable")
    items = np.empty(n, np.uint32)
    items[order] = np.arange(n, dtype=np.uint32)
    return items
