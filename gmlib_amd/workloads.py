"""Synthetic workload builders (inputs of tests and bench.py; not the path).

cube_edges(tet): unique edges of a tet mesh, sorted (i < j), lexicographic.
edge_laplacian3(nver, edges, seed, shift): the 3x3-block vector Laplacian of
an edge graph as block CSR (rows sorted, diagonal included):
   off-diagonal block (i, j) = -M_e,  M_e = w_e (I + 0.25 s_e s_e^T)
   diagonal block (i, i)     = sum of M_e over the edges at i + shift I
with w_e ~ U[0.5, 2] and s_e a random unit vector, both seeded: a sum of
PSD edge terms plus shift I, hence SPD.
"""
import numpy as np


def cube_edges(tet):
    tet = np.asarray(tet, np.int64)
    pairs = np.concatenate([tet[:, [a, b]] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))])
    pairs.sort(axis=1)
    key = np.unique(pairs[:, 0] << 32 | pairs[:, 1])
    return np.stack([key >> 32, key & 0xffffffff], axis=1).astype(np.int32)


def edge_laplacian3(nver, edges, seed=1, shift=0.01):
    rng = np.random.default_rng(seed)
    E = len(edges)
    w = rng.uniform(0.5, 2.0, E)
    s = rng.normal(size=(E, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    M = w[:, None, None] * (np.eye(3)[None] + 0.25 * s[:, :, None] * s[:, None, :])
    i, j = edges[:, 0].astype(np.int64), edges[:, 1].astype(np.int64)
    # entries: (i,j,-M), (j,i,-M), diagonals
    rows = np.concatenate([i, j, np.arange(nver)])
    cols = np.concatenate([j, i, np.arange(nver)])
    diag = np.zeros((nver, 3, 3))
    np.add.at(diag, i, M)
    np.add.at(diag, j, M)
    diag += shift * np.eye(3)[None]
    vals = np.concatenate([-M, -M, diag])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(nver + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, cols.astype(np.int32), np.ascontiguousarray(vals, np.float32)
