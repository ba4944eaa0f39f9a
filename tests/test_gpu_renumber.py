"""GPU Hilbert renumbering (GmlxHilbertRenumber).

A cube mesh numbered lexicographically and then shuffled is renumbered; the
result must be the same mesh (VerOld / EleOld are permutations, coordinates,
refs and element tables are exactly the permuted and renumbered input, the
tet qualities are the same multiset), the vertex and element Hilbert keys --
restated here on the host from the same definition (21-bit quantization of
the bounding box in FP32, Skilling's transform) -- must be non-decreasing,
and the edge index spread must fall to that of a Hilbert-ordered mesh."""
import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
BITS = 21


def hilbert_key(x, y, z, bits=BITS):
    X = [int(x), int(y), int(z)]
    M = 1 << (bits - 1)
    Q = M
    while Q > 1:
        P = Q - 1
        for i in range(3):
            if X[i] & Q:
                X[0] ^= P
            else:
                t = (X[0] ^ X[i]) & P
                X[0] ^= t
                X[i] ^= t
        Q >>= 1
    for i in range(1, 3):
        X[i] ^= X[i - 1]
    t = 0
    Q = M
    while Q > 1:
        if X[2] & Q:
            t ^= Q - 1
        Q >>= 1
    for i in range(3):
        X[i] ^= t
    h = 0
    for b in range(bits - 1, -1, -1):
        for i in range(3):
            h = (h << 1) | ((X[i] >> b) & 1)
    return h


def keys(pts, lo, sc):
    top = np.float32((1 << BITS) - 1)
    q = (pts[:, :3].astype(np.float32) - lo) * sc
    q = np.minimum(np.maximum(q, np.float32(0)), top).astype(np.uint32)
    return [hilbert_key(*r) for r in q]


def spread(tet):
    e = np.concatenate([tet[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    return float(np.abs(e[:, 0] - e[:, 1]).mean())


def test_hilbert_renumber_shuffled_cube():
    crd, ref, tet = G.cube_mesh(11, 0.2, 3, hilbert=False)
    V, T = len(crd), len(tet)
    rng = np.random.default_rng(0)
    pv, pt = rng.permutation(V), rng.permutation(T)
    inv = np.empty(V, np.int64)
    inv[pv] = np.arange(V)
    c0 = np.ascontiguousarray(crd[pv])
    r0 = np.ascontiguousarray(ref[pv])
    t0 = np.ascontiguousarray(inv[tet[pt]].astype(np.int32))
    tr0 = np.arange(T, dtype=np.int32) * 3
    c, r, t, tr = c0.copy(), r0.copy(), t0.copy(), tr0.copy()
    vold = np.empty(V, np.int32)
    eold = np.empty(T, np.int32)
    assert G.lib().GmlxHilbertRenumber(V, G.ptr(c), G.ptr(r), T, 4, G.ptr(t), G.ptr(tr), G.ptr(vold), G.ptr(eold), 0) == 1
    # same mesh, exactly permuted and renumbered
    assert np.array_equal(np.sort(vold), np.arange(V)) and np.array_equal(np.sort(eold), np.arange(T))
    vinv = np.empty(V, np.int64)
    vinv[vold] = np.arange(V)
    assert np.array_equal(c, c0[vold]) and np.array_equal(r, r0[vold])
    assert np.array_equal(t, vinv[t0[eold]]) and np.array_equal(tr, tr0[eold])
    q0 = np.sort(O.tet_quality64(t0, c0))
    q1 = np.sort(O.tet_quality64(t, c))
    assert np.array_equal(q0, q1)
    # Hilbert order of vertices and element barycentres (host restatement)
    lo = c0[:, :3].min(axis=0).astype(np.float32)
    hi = c0[:, :3].max(axis=0).astype(np.float32)
    sc = (np.float32((1 << BITS) - 1) / (hi - lo)).astype(np.float32)
    kv = keys(c, lo, sc)
    assert all(a <= b for a, b in zip(kv, kv[1:]))
    bary = np.zeros((T, 4), np.float32)
    for j in range(4):
        bary += c[t[:, j]]
    bary *= np.float32(0.25)
    ke = keys(bary, lo, sc)
    assert all(a <= b for a, b in zip(ke, ke[1:]))
    # locality: the edge index spread drops from the shuffle's to a Hilbert mesh's
    ref_spread = spread(G.cube_mesh(11, 0.2, 3, hilbert=True)[2])
    print("edge index spread: shuffled %.0f, renumbered %.0f, generated Hilbert mesh %.0f"
          % (spread(t0), spread(t), ref_spread))
    assert spread(t) < 0.25 * spread(t0) and spread(t) < 1.5 * ref_spread


def test_hilbert_renumber_is_deterministic_and_checks_input():
    crd, ref, tet = G.cube_mesh(5, 0.2, 1, hilbert=False)
    outs = []
    for _ in range(2):
        c, t = crd.copy(), tet.copy()
        assert G.lib().GmlxHilbertRenumber(len(c), G.ptr(c), None, len(t), 4, G.ptr(t), None, None, None, 0) == 1
        outs.append((c, t))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    bad = tet.copy()
    bad[0, 0] = len(crd)
    c = crd.copy()
    assert G.lib().GmlxHilbertRenumber(len(c), G.ptr(c), None, len(bad), 4, G.ptr(bad), None, None, None, 0) == 0


def test_hilbert_renumber_degree_classes():
    # Kuhn cube: interior vertices are in 24 tets, boundary ones in fewer; with
    # BasDeg 20 every vertex above 20 goes after the others, each class in
    # Hilbert order, and the mesh is the same
    crd, ref, tet = G.cube_mesh(7, 0.2, 5, hilbert=False)
    V, T = len(crd), len(tet)
    deg0 = np.bincount(tet.ravel(), minlength=V)
    c, t = crd.copy(), tet.copy()
    vold = np.empty(V, np.int32)
    assert G.lib().GmlxHilbertRenumber(V, G.ptr(c), None, T, 4, G.ptr(t), None, G.ptr(vold), None, 20) == 1
    deg = np.bincount(t.ravel(), minlength=V)
    assert np.array_equal(deg, deg0[vold]) and np.array_equal(c, crd[vold])
    high = deg > 20
    nlow = int((~high).sum())
    assert nlow > 0 and high[nlow:].all() and not high[:nlow].any()
    lo = crd[:, :3].min(axis=0).astype(np.float32)
    hi = crd[:, :3].max(axis=0).astype(np.float32)
    sc = (np.float32((1 << BITS) - 1) / (hi - lo)).astype(np.float32)
    kv = keys(c, lo, sc)
    assert all(a <= b for a, b in zip(kv[:nlow], kv[1:nlow])) and all(a <= b for a, b in zip(kv[nlow:], kv[nlow + 1:]))
    assert sorted(O.tet_quality64(t, c)) == sorted(O.tet_quality64(tet, crd))
