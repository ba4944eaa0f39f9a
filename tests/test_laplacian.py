"""GmlxEdgeLaplacian3 (host matrix builder of the block-CG jobs) against a
numpy restatement of the same definition: identical structure and
bit-identical FP32 blocks on small meshes; symmetric, positive definite
(smallest eigenvalue >= shift), and the block rows of a dense check sum to
shift I times the constant vector (the Laplacian part annihilates
constants)."""
import numpy as np
import pytest

import gmlib_amd as G

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z):
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return z ^ (z >> np.uint64(31))


def unit(z):
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def edge_blocks(seed, i, j):
    """M_e of edges (i, j), i < j, as [E, 3, 3] float64."""
    i = i.astype(np.uint64)
    j = j.astype(np.uint64)
    k = mix64(np.uint64(seed) ^ mix64((i << np.uint64(32)) | j))
    w = 0.5 + 1.5 * unit(mix64(k + np.uint64(1)))
    s = np.stack([2.0 * unit(mix64(k + np.uint64(2 + c))) - 1.0 for c in range(3)], axis=1)
    n = np.sqrt((s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2])
    small = n < 1e-3
    s[small] = [1.0, 0.0, 0.0]
    n[small] = 1.0
    s = s / n[:, None]
    eye = np.eye(3)
    return w[:, None, None] * (eye[None] + (0.25 * s[:, :, None]) * s[:, None, :])


def numpy_laplacian(nver, tet, seed, shift):
    pairs = np.concatenate([tet[:, [a, b]] for a in range(4) for b in range(4) if a != b])
    pairs = np.unique(pairs.astype(np.int64), axis=0)          # (row, col), row != col, sorted
    diag = np.stack([np.arange(nver), np.arange(nver)], axis=1)
    ent = np.unique(np.concatenate([pairs, diag]), axis=0)
    rows, cols = ent[:, 0], ent[:, 1]
    rowptr = np.zeros(nver + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    off = rows != cols
    Me = edge_blocks(seed, np.minimum(rows[off], cols[off]), np.maximum(rows[off], cols[off]))
    val = np.zeros((len(ent), 3, 3))
    val[off] = -Me
    D = np.zeros((nver, 3, 3))
    r_off = rows[off]
    for e in range(len(r_off)):                                 # ascending column order per row
        D[r_off[e]] += Me[e]
    D += shift * np.eye(3)[None]
    val[~off] = D
    return rowptr.astype(np.int32), cols.astype(np.int32), val.astype(np.float32)


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 11), (3, 12345)])
def test_edge_laplacian_matches_numpy(n, seed):
    crd, ref, tet = G.cube_mesh(n, 0.2, 5)
    rp, col, val = G.edge_laplacian3(len(crd), tet, seed, 0.01)
    rp0, col0, val0 = numpy_laplacian(len(crd), tet, seed, 0.01)
    assert np.array_equal(rp, rp0) and np.array_equal(col, col0)
    assert np.array_equal(val, val0)


def test_edge_laplacian_spd_and_nullspace():
    crd, ref, tet = G.cube_mesh(3, 0.2, 5)
    nv = len(crd)
    rp, col, val = G.edge_laplacian3(nv, tet, 7, 0.01)
    A = np.zeros((nv * 3, nv * 3))
    for r in range(nv):
        for k in range(rp[r], rp[r + 1]):
            A[r * 3:r * 3 + 3, col[k] * 3:col[k] * 3 + 3] = val[k]
    assert np.allclose(A, A.T, atol=0)                          # blocks of (i,j) and (j,i) are the same -M_e
    ev = np.linalg.eigvalsh(A)
    assert ev.min() >= 0.01 * (1 - 1e-4)
    one = np.tile(np.eye(3), (nv, 1))                           # constant vectors: A 1 = shift 1
    assert np.allclose(A @ one, 0.01 * one, atol=1e-5)


def test_edge_laplacian_rejects_bad_mesh():
    tet = np.array([[0, 1, 2, 5]], np.int32)
    rp = np.empty(5, np.int32)
    assert G.lib().GmlxEdgeLaplacian3(4, 1, G.ptr(tet), 1, 0.01, G.ptr(rp), None, None, 0) == -1
