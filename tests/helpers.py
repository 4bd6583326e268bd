"""Shared input builders for the test suite (seeded, deterministic)."""
import numpy as np

from nbx.disk import make_disk


def disk(n, mass=1e-5, r_min=0.5, r_max=4.0, seed=1, **kw):
    return make_disk(n, mass, r_min, r_max, seed=seed, **kw)


def clustered(n_clusters, per, spread, seed=0, box=3.0, mass=1e-9):
    """Tight clusters (for collisions / FoF) scattered in a disk-sized box."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-box, box, (n_clusters, 3))
    centers[:, 1] *= 0.01
    pts = (centers[:, None, :] + rng.normal(0, spread, (n_clusters, per, 3))).reshape(-1, 3)
    n = len(pts)
    v = rng.normal(0, 0.05, (n, 3))
    f = np.float32
    return dict(x=pts[:, 0].astype(f), y=pts[:, 1].astype(f), z=pts[:, 2].astype(f),
                vx=v[:, 0].astype(f), vy=v[:, 1].astype(f), vz=v[:, 2].astype(f),
                ax=np.zeros(n, f), ay=np.zeros(n, f), az=np.zeros(n, f),
                mass=np.full(n, mass, f), id=np.arange(n, dtype=np.int32))


def rel_err(a, b):
    """Per-row |a-b| / |b| (b the reference)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
