"""Synthetic point sets of the benchmark configurations (float32, seeded).

Each generator is deterministic for its seed; point order matters because
point indices break ties downstream.
"""
from __future__ import annotations

import numpy as np


def blobs_small(n: int = 10_000, seed: int = 0) -> np.ndarray:
    """Isotropic Gaussian mixture: 10 centres uniform in [-10, 10]^2, std 1.0
    (sklearn.datasets.make_blobs(n, 2, centers=10, cluster_std=1.0,
    center_box=(-10, 10), random_state=seed))."""
    from sklearn.datasets import make_blobs

    X = make_blobs(n_samples=n, n_features=2, centers=10, cluster_std=1.0, center_box=(-10.0, 10.0),
                   random_state=seed)[0]
    return X.astype(np.float32)


def anisotropic_mixture(n: int, d: int, n_clusters: int, noise_frac: float, seed: int,
                        center_range: float = 100.0, std_lo: float = 0.1, std_hi: float = 2.0,
                        zipf_s: float = 1.1) -> np.ndarray:
    """Embedding-like mixture: n_clusters anisotropic Gaussians (random
    rotation, per-axis std log-uniform in [std_lo, std_hi], Zipf(s) sizes)
    plus noise_frac uniform background over the clusters' bounding box,
    shuffled.  d=2, 50 clusters, 10% noise, seed 1 at n=2,000,000 is the
    2-D scale workload; d=3, 200 clusters, 5% noise, seed 2 at 10,000,000
    the 3-D one."""
    rng = np.random.default_rng(seed)
    n_noise = int(round(n * noise_frac))
    n_clu = n - n_noise
    r = np.arange(1, n_clusters + 1, dtype=np.float64)
    w = r ** (-zipf_s)
    sizes = rng.multinomial(n_clu, w / w.sum())
    centers = rng.uniform(-center_range, center_range, size=(n_clusters, d))
    parts = []
    for c in range(n_clusters):
        if d == 2:
            th = rng.uniform(0.0, 2.0 * np.pi)
            R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        else:
            Q, Rr = np.linalg.qr(rng.standard_normal((d, d)))
            R = Q * np.sign(np.diag(Rr))
        s = np.exp(rng.uniform(np.log(std_lo), np.log(std_hi), size=d))
        pts = centers[c] + (rng.standard_normal((sizes[c], d)) * s) @ R.T
        parts.append(pts)
    clu = np.concatenate(parts, axis=0)
    lo, hi = clu.min(axis=0), clu.max(axis=0)
    noise = rng.uniform(lo, hi, size=(n_noise, d))
    X = np.concatenate([clu, noise], axis=0)
    X = X[rng.permutation(X.shape[0])]
    return np.ascontiguousarray(X.astype(np.float32))


def scale_2d(n: int = 2_000_000, seed: int = 1) -> np.ndarray:
    """2-D scale workload: 50 anisotropic Gaussians + 10% noise."""
    return anisotropic_mixture(n, 2, 50, 0.10, seed)


def scale_3d(n: int = 10_000_000, seed: int = 2) -> np.ndarray:
    """3-D workload: 200 anisotropic Gaussians + 5% noise."""
    return anisotropic_mixture(n, 3, 200, 0.05, seed)


def gaussian_highdim(n: int = 1_000_000, d: int = 64, n_clusters: int = 100, seed: int = 3) -> np.ndarray:
    """100 equal Gaussian components, means ~ N(0, 4I), within ~ N(0, I)."""
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, 2.0, size=(n_clusters, d))
    lab = np.repeat(np.arange(n_clusters), int(np.ceil(n / n_clusters)))[:n]
    X = means[lab] + rng.standard_normal((n, d))
    X = X[rng.permutation(n)]
    return np.ascontiguousarray(X.astype(np.float32))


def batch_datasets(count: int = 4096, seed_base: int = 1000, n_lo: int = 500, n_hi: int = 50_000):
    """Independent small datasets: per dataset i (rng seed seed_base + i)
    n log-uniform in [n_lo, n_hi], d in {2, 3, 5, 8}, a 3-20 centre Gaussian
    mixture (centres uniform in [-10, 10]^d, std uniform 0.5-2.0) plus 5%
    uniform noise over its bounding box."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed_base + i)
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        d = int(rng.choice([2, 3, 5, 8]))
        k = int(rng.integers(3, 21))
        std = rng.uniform(0.5, 2.0)
        n_noise = int(0.05 * n)
        n_clu = n - n_noise
        centres = rng.uniform(-10.0, 10.0, size=(k, d))
        lab = rng.integers(0, k, n_clu)
        pts = centres[lab] + rng.normal(0.0, std, size=(n_clu, d))
        lo, hi = pts.min(0), pts.max(0)
        noise = rng.uniform(lo, hi, size=(n_noise, d))
        X = np.vstack([pts, noise])[rng.permutation(n)]
        out.append(np.ascontiguousarray(X, dtype=np.float32))
    return out
