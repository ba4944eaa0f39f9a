"""Edge and face shells, downlinks and neighbours of a host-only handle
against the pure-Python restatement of the reference's hash links
(tests/oracle_lib.py: GetItmNod / CalHshKey / AddHsh / GetHsh as used by
NewBallData, gmlib3.c:743-1168, and GmlSetNeighbours, 4178-4267), whose row
orders come from the chain walk itself.  Closed forms of the Kuhn cube and
the hex grid pin the restatement: every tet edge and face in as many shells
as its multiplicity, boundary faces in one shell, neighbour symmetry."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O
from tests.test_extract import get_block, hex_grid, set_block

TET, HEX, EDG, TRI, QAD = G.GmlTetrahedra, G.GmlHexahedra, G.GmlEdges, G.GmlTriangles, G.GmlQuadrilaterals


def link_rows(g, idx, n, w):
    """Lines of a link data item through GmlGetDataLine (w int pointers)."""
    out = np.zeros((n, w), np.int32)
    vals = [C.c_int() for _ in range(w)]
    for i in range(n):
        assert G.get_line(g.h, idx, i, *[C.byref(v) for v in vals])
        out[i] = [v.value for v in vals]
    return out


def table(g, typ, w):
    n = C.c_int(0)
    g.L.GmlGetMeshInfo(g.h, typ, C.byref(n), None)
    return get_block(g, typ, n.value, w)[0]


def tet_mesh(g, n):
    crd, ref, tet = G.cube_mesh(n, 0.1, 3)
    g.set_mesh(crd, ref, tet)
    assert g.L.GmlExtractEdges(g.h) > 0 and g.L.GmlExtractFaces(g.h) > 0
    return tet, table(g, EDG, 2), table(g, TRI, 3)


def hex_mesh(g, n):
    h = hex_grid(n)
    nv = (n + 1) ** 3
    assert g.L.GmlNewMeshData(g.h, G.GmlVertices, nv)
    set_block(g, HEX, h, np.zeros(len(h), np.int32))
    assert g.L.GmlExtractEdges(g.h) > 0 and g.L.GmlExtractFaces(g.h) > 0
    return h, table(g, EDG, 2), table(g, QAD, 4)


def shell_check(g, src, sk, dst, dk):
    assert g.L.GmlxBuildLink(g.h, sk, dk)
    want = O.ref_shells(src, sk, dst, dk)
    md = max(len(r) for r in want)
    deg, ball = g.uplink_table(sk, dk, len(src), max(md, 1))
    assert np.array_equal(deg, [len(r) for r in want])
    for s, r in enumerate(want):
        assert list(ball[s, :len(r)]) == r, s
    return deg


def down_check(g, src, sk, dst, dk):
    assert g.L.GmlxBuildLink(g.h, sk, dk)
    idx = g.L.GmlxGetLinkData(g.h, sk, dk)
    assert idx
    n, w, _, _ = g.link_info(sk, dk)
    got = link_rows(g, idx, n, w)
    assert np.array_equal(got, O.ref_downlink(src, sk, dst, dk, w))
    return got


def ngb_check(g, tab, kind):
    idx = g.L.GmlSetNeighbours(g.h, kind)
    assert idx and g.L.GmlxGetLinkData(g.h, kind, kind) == idx
    n, w, _, _ = g.link_info(kind, kind)
    got = link_rows(g, idx, n, w)
    assert np.array_equal(got, O.ref_neighbours(tab, kind, w))
    return got


@pytest.mark.parametrize("n", [1, 2])
def test_tet_shells_downlinks_neighbours(n):
    with G.Gml(host_only=True) as g:
        tet, edg, tri = tet_mesh(g, n)
        T = len(tet)
        deg = shell_check(g, edg, EDG, tet, TET)
        assert deg.sum() == 6 * T
        deg = shell_check(g, tri, TRI, tet, TET)
        assert deg.sum() == 4 * T and (deg == 1).sum() == 12 * n * n and deg.max() == 2
        d = down_check(g, tet, TET, edg, EDG)
        for j, (a, b) in enumerate(O.EDGE_LOC[TET]):     # each entry is the edge with those nodes
            assert np.array_equal(edg[d[:, j]], np.sort(tet[:, [a, b]], axis=1))
        d = down_check(g, tet, TET, tri, TRI)
        for j, f in enumerate(O.FACE_LOC[TET]):
            assert np.array_equal(tri[d[:, j]], np.sort(tet[:, list(f)], axis=1))
        ngb = ngb_check(g, tet, TET)
        # symmetry: a neighbour across a face lists the element back
        for i in range(T):
            for j in range(4):
                k = ngb[i, j]
                if k:
                    assert i in ngb[k, :4] or (i == 0)


def test_triangle_and_edge_neighbours_many_sharing():
    # the triangle table of all tet faces: interior edges are shared by many
    # triangles, so the lowest other element is the rule that decides
    with G.Gml(host_only=True) as g:
        tet, edg, tri = tet_mesh(g, 2)
        ngb_check(g, tri, TRI)
        ngb_check(g, edg, EDG)
        shell_check(g, edg, EDG, tri, TRI)
        down_check(g, tri, TRI, edg, EDG)


def test_hex_shells_downlinks_neighbours():
    n = 3
    with G.Gml(host_only=True) as g:
        h, edg, qad = hex_mesh(g, n)
        deg = shell_check(g, qad, QAD, h, HEX)
        assert deg.sum() == 6 * n ** 3 and (deg == 1).sum() == 6 * n * n
        shell_check(g, edg, EDG, h, HEX)
        down_check(g, h, HEX, qad, QAD)
        down_check(g, h, HEX, edg, EDG)
        ngb = ngb_check(g, h, HEX)
        assert (ngb[1:, :6] == 0).sum() + (ngb[0, :6] == 0).sum() >= 6 * n * n


def test_neighbours_refused_for_prisms_and_twice():
    with G.Gml(host_only=True) as g:
        tet_mesh(g, 1)
        assert g.L.GmlSetNeighbours(g.h, G.GmlPrisms) == 0
        assert g.L.GmlSetNeighbours(g.h, TET) > 0
        assert g.L.GmlSetNeighbours(g.h, TET) == 0
