"""Shared helpers of the test suite: seeded meshes, layout conversions."""
import numpy as np


def box_vertices(dim, nc, perturb=0.0, seed=0, lo=0.0, hi=1.0):
    """Vertices of a structured box mesh with nc cells per direction
    (x fastest).  Interior vertices are moved by a seeded uniform
    perturbation of up to `perturb` times the cell size in every coordinate;
    boundary vertices stay on the boundary."""
    n = [nc[d] + 1 for d in range(dim)]
    axes = [np.linspace(lo, hi, n[d]) for d in range(dim)]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    X = np.stack([g.reshape(-1) for g in grids], axis=1)
    if perturb > 0:
        rng = np.random.default_rng(seed)
        h = [(hi - lo) / nc[d] for d in range(dim)]
        idx = np.indices(n[::-1]).reshape(dim, -1)[::-1]
        interior = np.ones(X.shape[0], dtype=bool)
        for d in range(dim):
            interior &= (idx[d] > 0) & (idx[d] < n[d] - 1)
        for d in range(dim):
            X[interior, d] += perturb * h[d] * rng.uniform(-1, 1, interior.sum())
    return np.ascontiguousarray(X)


def log_uniform(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def blocks_to_nodes(x, nt):
    return np.ascontiguousarray(np.asarray(x).reshape(nt, -1).T.reshape(-1))


def nodes_to_blocks(x, nt):
    return np.ascontiguousarray(np.asarray(x).reshape(-1, nt).T.reshape(-1))


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
