"""GmlExtractEdges / GmlExtractFaces (host, no GPU) against a brute-force
restatement of the reference's rule (sources/gmlib3.c:3719-4171): the items
of an existing table first, in order, with their refs; then the items of the
element kinds above, element by element and local item by local item
(ItmEdgNod / ItmFacNod order), each at its first appearance, ref 0; nodes
stored sorted.  Counts are also checked against the closed forms of the
Kuhn cube."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G

TET_EDGES = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 1), (3, 2)]          # ItmEdgNod, tets
TET_FACES = [(1, 2, 3), (2, 0, 3), (3, 0, 1), (0, 2, 1)]            # ItmFacNod, tets


def first_appearance(existing, tet, local):
    seen, items, refs = set(), [], []
    for nodes, r in existing:
        k = tuple(sorted(nodes))
        if k not in seen:
            seen.add(k)
            items.append(k)
            refs.append(r)
    for t in tet:
        for loc in local:
            k = tuple(sorted(int(t[i]) for i in loc))
            if k not in seen:
                seen.add(k)
                items.append(k)
                refs.append(0)
    return np.array(items, np.int32), np.array(refs, np.int32)


def get_block(g, typ, n, w):
    e = np.empty((n, w), np.int32)
    r = np.empty(n, np.int32)
    assert g.L.GmlGetDataBlock(g.h, typ, 0, n - 1, G.ptr(e), C.c_void_p(e.ctypes.data + (n - 1) * 4 * w),
                               G.ptr(r), C.c_void_p(r.ctypes.data + (n - 1) * 4))
    return e, r


def set_block(g, typ, nodes, refs):
    n, w = nodes.shape
    idx = g.L.GmlNewMeshData(g.h, typ, n)
    assert idx
    assert g.L.GmlSetDataBlock(g.h, typ, 0, n - 1, G.ptr(nodes), C.c_void_p(nodes.ctypes.data + (n - 1) * 4 * w),
                               G.ptr(refs), C.c_void_p(refs.ctypes.data + (n - 1) * 4))


@pytest.mark.parametrize("n", [1, 3])
def test_extract_edges_order_and_refs(n):
    crd, ref, tet = G.cube_mesh(n, 0.1, 2)
    old = np.array([[tet[0, 1], tet[0, 0]], [0, len(crd) - 1], [tet[5, 2], tet[5, 3]]], np.int32)
    old_ref = np.array([5, 6, 7], np.int32)
    with G.Gml(host_only=True) as g:
        g.set_mesh(crd, ref, tet)
        set_block(g, G.GmlEdges, old, old_ref)
        ne = g.L.GmlExtractEdges(g.h)
        e, r = get_block(g, G.GmlEdges, ne, 2)
    want, wref = first_appearance(zip(old, old_ref), tet, TET_EDGES)
    assert ne == len(want)
    assert np.array_equal(e, want) and np.array_equal(r, wref)
    # closed form of the Kuhn cube, plus the given edges that are not mesh edges
    mesh = {tuple(sorted((int(t[a]), int(t[b])))) for t in tet for a, b in TET_EDGES}
    extra = len({tuple(sorted(map(int, x))) for x in old} - mesh)
    assert ne - extra == 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3


@pytest.mark.parametrize("n", [1, 3])
def test_extract_faces_order_and_refs(n):
    crd, ref, tet = G.cube_mesh(n, 0.1, 4)
    surf = np.array([sorted(tet[2, list(TET_FACES[1])])[::-1], [0, 1, len(crd) - 1]], np.int32)
    surf_ref = np.array([11, 12], np.int32)
    with G.Gml(host_only=True) as g:
        g.set_mesh(crd, ref, tet)
        set_block(g, G.GmlTriangles, surf, surf_ref)
        nf = g.L.GmlExtractFaces(g.h)
        f, r = get_block(g, G.GmlTriangles, nf, 3)
    want, wref = first_appearance(zip(surf, surf_ref), tet, TET_FACES)
    assert nf == len(want)
    assert np.array_equal(f, want) and np.array_equal(r, wref)
    # every face once: 4T = 2 * inner + boundary, boundary = 12 n^2, plus the given
    # triangles that are not mesh faces
    T = len(tet)
    mesh = {tuple(sorted(int(t[i]) for i in loc)) for t in tet for loc in TET_FACES}
    extra = len({tuple(sorted(map(int, x))) for x in surf} - mesh)
    assert nf - extra == (4 * T + 12 * n * n) // 2


def test_extract_refused_while_links_exist():
    crd, ref, tet = G.cube_mesh(2, 0.1, 1)
    with G.Gml(host_only=True) as g:
        g.set_mesh(crd, ref, tet)
        assert g.L.GmlExtractEdges(g.h) > 0
        assert g.L.GmlxBuildLink(g.h, G.GmlEdges, G.GmlTetrahedra)       # an edge shell now exists
        assert g.L.GmlExtractEdges(g.h) == 0
