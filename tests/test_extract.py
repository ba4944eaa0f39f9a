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


# --- the oracle's restatement of the reference hash (orc_extract_items) ------

from tests import oracle_lib as O  # noqa: E402


def hex_grid(n):
    """Hexes of an n^3 grid, GMlib node order (bottom 0-3, top 4-7)."""
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = (x.transpose(2, 1, 0).ravel() for x in (i, j, k))
    v = lambda a, b, c: (c * (n + 1) + b) * (n + 1) + a
    return np.stack([v(i, j, k), v(i + 1, j, k), v(i + 1, j + 1, k), v(i, j + 1, k),
                     v(i, j, k + 1), v(i + 1, j, k + 1), v(i + 1, j + 1, k + 1), v(i, j + 1, k + 1)], 1).astype(np.int32)


def brute(deg, tables, old=(), old_ref=()):
    """first_appearance over several element kinds (oracle_lib's local tables)."""
    seen, items, refs = set(), [], []
    for nodes, r in zip(old, old_ref):
        k = tuple(sorted(int(x) for x in nodes))
        if k not in seen:
            seen.add(k); items.append(k); refs.append(int(r))
    for kind, t in tables:
        loc = O.EDGE_LOC[kind] if deg == 2 else [f for f in O.FACE_LOC[kind] if len(f) == deg]
        for e in t:
            for f in loc:
                k = tuple(sorted(int(e[i]) for i in f))
                if k not in seen:
                    seen.add(k); items.append(k); refs.append(0)
    return np.array(items, np.int32).reshape(-1, deg), np.array(refs, np.int32)


@pytest.mark.parametrize("tabsiz", [0, 7])
def test_oracle_extract_matches_brute_force(tabsiz):
    # tabsiz 7: every chain long, the node comparison along chains decides
    crd, ref, tet = G.cube_mesh(3, 0.1, 5)
    old = np.array([[tet[4, 1], tet[4, 3]], [2, 9], [tet[4, 3], tet[4, 1]]], np.int32)
    n, r = O.extract_items(2, [(G.GmlTetrahedra, tet)], old, np.array([3, 4, 5], np.int32), tabsiz)
    w, wr = brute(2, [(G.GmlTetrahedra, tet)], old, [3, 4, 5])
    assert np.array_equal(n, w) and np.array_equal(r, wr)
    n, r = O.extract_items(3, [(G.GmlTetrahedra, tet)], tabsiz=tabsiz)
    w, wr = brute(3, [(G.GmlTetrahedra, tet)])
    assert np.array_equal(n, w) and np.array_equal(r, wr)


def test_oracle_extract_hex_grid_closed_forms():
    # n^3 hexes: 3 n (n+1)^2 edges and 3 n^2 (n+1) quads; no triangles
    n = 4
    h = hex_grid(n)
    e, _ = O.extract_items(2, [(G.GmlHexahedra, h)])
    q, _ = O.extract_items(4, [(G.GmlHexahedra, h)])
    t, _ = O.extract_items(3, [(G.GmlHexahedra, h)])
    assert len(e) == 3 * n * (n + 1) ** 2 and len(q) == 3 * n * n * (n + 1) and len(t) == 0
    # hexes split into two prisms each: one diagonal edge and two triangles more per cell,
    # the diagonal quad faces shared by the two prisms of a cell
    pr = np.concatenate([h[:, [0, 1, 2, 4, 5, 6]], h[:, [0, 2, 3, 4, 6, 7]]])
    e, _ = O.extract_items(2, [(G.GmlPrisms, pr)])
    t, _ = O.extract_items(3, [(G.GmlPrisms, pr)])
    q, _ = O.extract_items(4, [(G.GmlPrisms, pr)])
    assert len(e) == 3 * n * (n + 1) ** 2 + n * n * (n + 1)        # a diagonal per z-face
    assert len(t) == 2 * n * n * (n + 1)                           # two triangles per z-face
    assert len(q) == 2 * n * n * (n + 1) + n ** 3                  # x, y faces + one diagonal quad per cell
    # a pyramid (bottom quad 0-3, apex 4) and its 8 edges, 4 triangles, 1 quad
    py = np.array([[0, 1, 2, 3, 4]], np.int32)
    assert len(O.extract_items(2, [(G.GmlPyramids, py)])[0]) == 8
    assert len(O.extract_items(3, [(G.GmlPyramids, py)])[0]) == 4
    assert len(O.extract_items(4, [(G.GmlPyramids, py)])[0]) == 1


def test_host_extract_mixed_kinds_matches_oracle():
    n = 3
    h = hex_grid(n)
    nv = (n + 1) ** 3
    kinds = {G.GmlHexahedra: h[0::3], G.GmlPrisms: np.concatenate([h[1::3][:, [0, 1, 2, 4, 5, 6]],
                                                                     h[1::3][:, [0, 2, 3, 4, 6, 7]]]),
             G.GmlPyramids: h[2::3][:, :5], G.GmlTetrahedra: h[2::3][:, [1, 2, 4, 5]]}
    quads = np.array([h[0, [7, 6, 5, 4]], [0, 1, 2, 3]], np.int32)
    crd = np.zeros((nv, 4), np.float32)
    for what in (2, 3):
        with G.Gml(host_only=True) as g:
            g.L.GmlNewMeshData(g.h, G.GmlVertices, nv)
            for typ, t in kinds.items():
                set_block(g, typ, np.ascontiguousarray(t, np.int32), np.zeros(len(t), np.int32))
            set_block(g, G.GmlQuadrilaterals, quads, np.array([31, 32], np.int32))
            if what == 2:
                ne = g.L.GmlExtractEdges(g.h)
                got = {2: get_block(g, G.GmlEdges, ne, 2)}
            else:
                nf = g.L.GmlExtractFaces(g.h)
                cnt = [C.c_int(0), C.c_int(0)]
                g.L.GmlGetMeshInfo(g.h, G.GmlTriangles, C.byref(cnt[0]), None)
                g.L.GmlGetMeshInfo(g.h, G.GmlQuadrilaterals, C.byref(cnt[1]), None)
                assert nf == cnt[0].value + cnt[1].value
                got = {3: get_block(g, G.GmlTriangles, cnt[0].value, 3), 4: get_block(g, G.GmlQuadrilaterals, cnt[1].value, 4)}
        tabs = [(k, np.ascontiguousarray(t, np.int32)) for k, t in sorted(kinds.items())]
        for deg, (nodes, refs) in got.items():
            if deg == 2:
                want = O.extract_items(2, [(G.GmlQuadrilaterals, quads)] + tabs)
            elif deg == 3:
                want = O.extract_items(3, tabs)
            else:
                want = O.extract_items(4, tabs, quads, np.array([31, 32], np.int32))
            assert np.array_equal(nodes, want[0]) and np.array_equal(refs, want[1]), deg
