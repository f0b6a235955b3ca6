"""Test helpers: an independent restatement of the reference's structured
Kuhn cube generator (tests/mesh_utils.py:16-56 of the reference) and
small seeded fields."""

import numpy as np


def structured_cube_mesh(n: int, L: float = 1.0):
    lin = np.linspace(0.0, L, n + 1)
    X, Y, Z = np.meshgrid(lin, lin, lin, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float64)

    def idx(i, j, k):
        return i * (n + 1) * (n + 1) + j * (n + 1) + k

    unit = {0: (1, 0, 0), 1: (0, 1, 0), 2: (0, 0, 1)}
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    el = np.empty((6 * n ** 3, 4), dtype=np.int32)
    e = 0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for p in perms:
                    v = [np.array([i, j, k])]
                    for a in p:
                        v.append(v[-1] + np.array(unit[a]))
                    el[e] = [idx(*x) for x in v]
                    e += 1
    return nodes, el


def cube_boundary_mask(nodes, L=1.0, tol=1e-9):
    return np.any(nodes < tol, axis=1) | np.any(nodes > L - tol, axis=1)


def jittered_cube(n, amp=0.15, seed=0):
    nodes, el = structured_cube_mesh(n)
    b = cube_boundary_mask(nodes)
    rng = np.random.default_rng(seed)
    nodes = nodes.copy()
    nodes[~b] += rng.uniform(-amp, amp, size=(int((~b).sum()), 3)) / n
    return nodes, el


def transonic_field(nodes, seed=42, noise=1e-3):
    """phi = 1.25 x + 0.35 sin(1.7 x) + noise (the reference's constructed
    multi-regime field, cases/demo/p7_diff_flux/run_demo.py:200-203)."""
    rng = np.random.default_rng(seed)
    x = nodes[:, 0]
    return 1.25 * x + 0.35 * np.sin(1.7 * x) + noise * rng.standard_normal(len(x))
