"""GmlExtractEdges / GmlExtractFaces on a device handle (a mesh of tets only
through the vertex -> tet ball; otherwise packed node keys, deduplicated per
work-group, then bucketed by lowest node or sorted) against the same calls on
a host-only handle
(the ordered-map form, itself checked against a brute-force restatement of
sources/gmlib3.c:3719-4171 in test_extract.py) and against that restatement
directly: bit-identical tables, refs and counts.  Covers the one-word key
(b bits per node in one 64-bit word) and the two-word key (the same b-bit
packing spilling into a high word, forced by a vertex count above 2^21 for
triangles and 2^16 for quads), mixed element kinds
(tets, pyramids, prisms, hexes), an existing table with refs and duplicates,
and the count closed forms of a large Kuhn cube."""
import ctypes as C
import time

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O
from tests.test_extract import TET_EDGES, TET_FACES, first_appearance, get_block, set_block

pytestmark = pytest.mark.gpu


def set_vertices(g, nv):
    crd = np.zeros((nv, 4), np.float32)
    crd[:, 0] = np.arange(nv, dtype=np.float32)
    ref = np.zeros(nv, np.int32)
    assert g.L.GmlNewMeshData(g.h, G.GmlVertices, nv)
    assert g.L.GmlSetDataBlock(g.h, G.GmlVertices, 0, nv - 1, G.ptr(crd), C.c_void_p(crd.ctypes.data + (nv - 1) * 16),
                               G.ptr(ref), C.c_void_p(ref.ctypes.data + (nv - 1) * 4))


def hex_grid(n):
    """Hexes of an n^3 grid, GMlib node order (bottom 0-3, top 4-7)."""
    idx = lambda i, j, k: (k * (n + 1) + j) * (n + 1) + i
    h = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                b = [idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k)]
                t = [idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)]
                h.append(b + t)
    return np.array(h, np.int32), (n + 1) ** 3


def mixed_mesh(n, spread=1):
    """(nv, {kind: table}): an n^3 hex grid, odd cells split into two prisms,
    every third even cell into a pyramid and a tet pair; node ids scaled by
    spread (unused vertices in between) to reach the two-word key form."""
    hexes, nv = hex_grid(n)
    out = {G.GmlHexahedra: [], G.GmlPrisms: [], G.GmlPyramids: [], G.GmlTetrahedra: []}
    for c, h in enumerate(hexes):
        if c % 2:
            out[G.GmlPrisms] += [[h[0], h[1], h[2], h[4], h[5], h[6]], [h[0], h[2], h[3], h[4], h[6], h[7]]]
        elif c % 3 == 0:
            out[G.GmlPyramids].append([h[0], h[1], h[2], h[3], h[4]])
            out[G.GmlTetrahedra] += [[h[1], h[2], h[4], h[5]], [h[2], h[3], h[4], h[7]]]
        else:
            out[G.GmlHexahedra].append(list(h))
    tabs = {k: np.array(v, np.int32) * spread for k, v in out.items() if v}
    return nv * spread, tabs


def build_mixed(nv, tabs, extra=None):
    def build(g):
        set_vertices(g, nv)
        for typ, t in tabs.items():
            set_block(g, typ, t, np.zeros(len(t), np.int32))
        if extra:
            for typ, (t, r) in extra.items():
                set_block(g, typ, t, r)
    return build


def run(host_only, build, what):
    """Extraction on one handle: count and {kind: (nodes, refs)}."""
    with G.Gml(host_only=host_only) as g:
        build(g)
        res = {}
        if what == "edges":
            n = g.L.GmlExtractEdges(g.h)
            res[G.GmlEdges] = get_block(g, G.GmlEdges, n, 2)
        else:
            n = g.L.GmlExtractFaces(g.h)
            for typ, w in ((G.GmlTriangles, 3), (G.GmlQuadrilaterals, 4)):
                m = count_of(g, typ)
                if m:
                    res[typ] = get_block(g, typ, m, w)
        return n, res


def count_of(g, typ):
    """Lines of a mesh table (0 if absent)."""
    n = C.c_int(0)
    g.L.GmlGetMeshInfo(g.h, typ, C.byref(n), None)
    return n.value


def same(a, b):
    assert a[0] == b[0]
    assert sorted(a[1]) == sorted(b[1])
    for k in a[1]:
        assert np.array_equal(a[1][k][0], b[1][k][0]), k
        assert np.array_equal(a[1][k][1], b[1][k][1]), k


@pytest.mark.parametrize("n", [1, 4])
def test_gpu_extract_tets_against_restatement(n):
    crd, ref, tet = G.cube_mesh(n, 0.1, 2)
    old = np.array([[tet[0, 1], tet[0, 0]], [0, len(crd) - 1], [tet[5, 2], tet[5, 3]], [tet[0, 0], tet[0, 1]]], np.int32)
    old_ref = np.array([5, 6, 7, 8], np.int32)
    surf = np.array([sorted(tet[2, list(TET_FACES[1])])[::-1], [0, 1, len(crd) - 1]], np.int32)
    surf_ref = np.array([11, 12], np.int32)

    def build_e(g):
        g.set_mesh(crd, ref, tet)
        set_block(g, G.GmlEdges, old, old_ref)

    def build_f(g):
        g.set_mesh(crd, ref, tet)
        set_block(g, G.GmlTriangles, surf, surf_ref)

    ne, e = run(False, build_e, "edges")
    want, wref = first_appearance(zip(old, old_ref), tet, TET_EDGES)
    assert ne == len(want)
    assert np.array_equal(e[G.GmlEdges][0], want) and np.array_equal(e[G.GmlEdges][1], wref)
    nf, f = run(False, build_f, "faces")
    want, wref = first_appearance(zip(surf, surf_ref), tet, TET_FACES)
    assert nf == len(want) and sorted(f) == [G.GmlTriangles]
    assert np.array_equal(f[G.GmlTriangles][0], want) and np.array_equal(f[G.GmlTriangles][1], wref)


@pytest.mark.parametrize("spread", [1, 800, 40000])
def test_gpu_extract_mixed_kinds_same_as_host(spread):
    # spread 800: 51,200 vertices (16-bit ids) -- quad keys fill all 64 bits of
    # one word (the round-based block pass), edges and triangles take the
    # compare-and-swap pass; spread 40000: 2.56M vertices (22-bit ids) --
    # two-word keys for triangles and quads
    nv, tabs = mixed_mesh(3, spread)
    quads = np.array([tabs[G.GmlHexahedra][0, [4, 5, 6, 7]][::-1], [0, spread, 2 * spread, 3 * spread]], np.int32)
    extra = {G.GmlQuadrilaterals: (quads, np.array([21, 22], np.int32))}
    build = build_mixed(nv, tabs, extra)
    srcs = sorted(tabs.items())
    qref = np.array([21, 22], np.int32)
    want = {G.GmlEdges: O.extract_items(2, [(G.GmlQuadrilaterals, quads)] + srcs),
            G.GmlTriangles: O.extract_items(3, srcs),
            G.GmlQuadrilaterals: O.extract_items(4, srcs, quads, qref)}
    for what in ("edges", "faces"):
        dev, host = run(False, build, what), run(True, build, what)
        same(dev, host)
        assert dev[0] > 0
        for typ, (nodes, refs) in dev[1].items():
            assert np.array_equal(nodes, want[typ][0]) and np.array_equal(refs, want[typ][1]), typ


def test_gpu_extract_large_cube_counts():
    n = 64
    crd, ref, tet = G.cube_mesh(n, 0.1, 3)
    T = len(tet)
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        t0 = time.perf_counter()
        ne = g.L.GmlExtractEdges(g.h)
        t1 = time.perf_counter()
        nf = g.L.GmlExtractFaces(g.h)
        t2 = time.perf_counter()
        e, _ = get_block(g, G.GmlEdges, ne, 2)
        f, _ = get_block(g, G.GmlTriangles, nf, 3)
    print("extraction of %d tets: %d edges in %.1f ms, %d faces in %.1f ms"
          % (T, ne, 1e3 * (t1 - t0), nf, 1e3 * (t2 - t1)))
    assert ne == 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    assert nf == (4 * T + 12 * n * n) // 2
    # sorted nodes, all distinct, every tet's items present
    assert (e[:, 0] < e[:, 1]).all() and (f[:, 0] < f[:, 1]).all() and (f[:, 1] < f[:, 2]).all()
    ke = e[:, 0].astype(np.int64) * len(crd) + e[:, 1]
    assert len(np.unique(ke)) == ne
    te = np.sort(tet[:, [0, 1]], axis=1)
    kt = te[:, 0].astype(np.int64) * len(crd) + te[:, 1]
    assert np.isin(kt, ke).all()
    # first appearance: the first tet's edges lead, in local order
    want, _ = first_appearance([], tet[:2], TET_EDGES)
    assert np.array_equal(e[:len(want)], want)


def test_gpu_extract_on_reference_sample_mesh(tmp_path):
    # the reference's 2D sample mesh (a NACA 0012 airfoil triangulation, its
    # 465 boundary edges with refs): GmlImportMesh, then the edge table of the
    # whole mesh, the boundary edges first with their refs
    import lzma
    import os
    data = lzma.open(os.path.join(os.path.dirname(__file__), "golden", "naca012_2D_triangles.meshb.xz")).read()
    path = str(tmp_path / "naca.meshb")
    open(path, "wb").write(data)
    res = []
    for host_only in (False, True):
        with G.Gml(host_only=host_only) as g:
            assert g.L.GmlImportMesh(C.c_size_t(g.h), C.c_char_p(path.encode()), C.c_int(4), C.c_int(5),
                                     C.c_int(6), C.c_int(0)) == 3
            old = get_block(g, G.GmlEdges, 465, 2)
            ne = g.L.GmlExtractEdges(g.h)
            res.append((ne, get_block(g, G.GmlEdges, ne, 2), old))
    (ne, (e, r), (old, oref)), (ne2, (e2, r2), _) = res
    assert ne == ne2 and np.array_equal(e, e2) and np.array_equal(r, r2)
    # Euler for a plane triangulation with one hole: V - E + T = 0
    assert ne == 36624 + 72783
    assert np.array_equal(e[:465], np.sort(old, axis=1)) and np.array_equal(r[:465], oref) and (r[465:] == 0).all()


def test_gpu_extract_small_and_empty_meshes():
    # one tet: its 6 edges and 4 faces in local order; vertices only: nothing
    crd = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
    tet = np.array([[0, 1, 2, 3]], np.int32)
    for host_only in (False, True):
        with G.Gml(host_only=host_only) as g:
            vi = g.L.GmlNewMeshData(g.h, G.GmlVertices, 4)
            ti = g.L.GmlNewMeshData(g.h, G.GmlTetrahedra, 1)
            assert vi and ti
            for i in range(4):
                assert G.set_line(g.h, vi, i, *[C.c_double(float(x)) for x in crd[i, :3]], C.c_int(0))
            assert G.set_line(g.h, ti, 0, *[C.c_int(int(x)) for x in tet[0]], C.c_int(0))
            ne = g.L.GmlExtractEdges(g.h)
            nf = g.L.GmlExtractFaces(g.h)
            e = get_block(g, G.GmlEdges, ne, 2)[0]
            f = get_block(g, G.GmlTriangles, nf, 3)[0]
        want_e, _ = first_appearance([], tet, TET_EDGES)
        want_f, _ = first_appearance([], tet, TET_FACES)
        assert ne == 6 and nf == 4 and np.array_equal(e, want_e) and np.array_equal(f, want_f), host_only
        with G.Gml(host_only=host_only) as g:
            assert g.L.GmlNewMeshData(g.h, G.GmlVertices, 4)
            assert g.L.GmlExtractEdges(g.h) == 0 and g.L.GmlExtractFaces(g.h) == 0
            assert not g.L.GmlGetMeshInfo(g.h, G.GmlEdges, None, None)
            assert not g.L.GmlGetMeshInfo(g.h, G.GmlTriangles, None, None)


def test_gpu_extract_reference_sample_mesh(tmp_path):
    # the reference's naca012 sample (36,624 vertices, 72,783 triangles and
    # 465 boundary edges with refs), read on a device handle: the GPU
    # extraction keeps the file's edges first with their refs, then the
    # triangles' other edges in first-appearance order, as the oracle's
    # restatement of the reference hash
    import lzma
    import os
    from tests.test_meshb import GMF, imp, mesh_tables
    data = lzma.open(os.path.join(os.path.dirname(__file__), "golden", "naca012_2D_triangles.meshb.xz")).read()
    path = str(tmp_path / "naca.meshb")
    open(path, "wb").write(data)
    with G.Gml() as g:
        assert imp(g, path, GMF["Vertices"], GMF["Edges"], GMF["Triangles"]) == 3
        t = mesh_tables(g)
        ne = g.L.GmlExtractEdges(g.h)
        got = mesh_tables(g)["Edges"]
    old, oref = t["Edges"]
    tri = t["Triangles"][0]
    want, wref = O.extract_items(2, [(G.GmlTriangles, tri)], old, oref)
    assert ne == len(want) and np.array_equal(got[0], want) and np.array_equal(got[1], wref)
    assert np.array_equal(got[0][:len(old)], np.sort(old, axis=1))          # the file's edges, nodes sorted
    assert (got[1][:len(old)] == oref).all() and (got[1][len(old):] == 0).all()


def extract_logged(g, what, capfd):
    """run one extraction with the debug log on: (count, path taken)"""
    capfd.readouterr()
    g.L.GmlDebugOn(g.h)
    n = g.L.GmlExtractEdges(g.h) if what == "edges" else g.L.GmlExtractFaces(g.h)
    g.L.GmlDebugOff(g.h)
    out = capfd.readouterr().out
    path = ("vertex-ball" if "vertex-ball path" in out else "bucketed" if "bucketed path" in out
            else "sorted" if "sorted path" in out else None)
    return n, path


@pytest.mark.parametrize("old", [False, True])
@pytest.mark.parametrize("shuffle", [False, True])
def test_gpu_extract_cube_tables_against_oracle(capfd, shuffle, old):
    # a 40^3 cube (384,000 tets; 2.3M edge items, 1.5M face items), in its
    # Hilbert numbering and with vertices and tets shuffled (buckets then
    # spread over the whole vertex range): tables and order bit-identical to
    # the oracle's restatement of the reference hash.  A mesh of tets only
    # takes the vertex-ball path; with a two-row table of the kind already
    # there (kept first, with its ref) the bucketed path runs
    crd, ref, tet = G.cube_mesh(40, 0.2, 17)
    if shuffle:
        rng = np.random.default_rng(3)
        pv = rng.permutation(len(crd))
        inv = np.empty_like(pv)
        inv[pv] = np.arange(len(crd))
        crd, ref = np.ascontiguousarray(crd[pv]), np.ascontiguousarray(ref[pv])
        tet = np.ascontiguousarray(inv[tet[rng.permutation(len(tet))]].astype(np.int32))
    oe = np.array([[tet[7, 2], tet[7, 0]], [tet[11, 1], tet[11, 3]]], np.int32) if old else None
    of = np.array([[tet[9, 3], tet[9, 1], tet[9, 2]], [0, 1, 2]], np.int32) if old else None
    oref = np.array([31, 32], np.int32) if old else None
    want_e = O.extract_items(2, [(G.GmlTetrahedra, tet)], oe, oref)
    want_f = O.extract_items(3, [(G.GmlTetrahedra, tet)], of, oref)
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        if old:
            set_block(g, G.GmlEdges, oe, oref)
        ne, pe = extract_logged(g, "edges", capfd)
        if old:
            set_block(g, G.GmlTriangles, of, oref)
        nf, pf = extract_logged(g, "faces", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        f = get_block(g, G.GmlTriangles, nf, 3)
    path = "bucketed" if old else "vertex-ball"
    assert pe == path and pf == path
    assert ne == len(want_e[0]) and nf == len(want_f[0])
    assert np.array_equal(e[0], want_e[0]) and np.array_equal(e[1], want_e[1])
    assert np.array_equal(f[0], want_f[0]) and np.array_equal(f[1], want_f[1])


def test_gpu_extract_large_bucket_takes_sorted_path(capfd):
    # a fan of 3000 tets (0, i, i+1, i+2): vertex 0 is the lowest node of
    # thousands of distinct edges and faces, more than a bucket takes, so the
    # call goes down the sorted path; its tables must still equal the oracle's
    nt = 3000
    tet = np.array([[0, i, i + 1, i + 2] for i in range(1, nt + 1)], np.int32)
    nv = nt + 3
    want_e = O.extract_items(2, [(G.GmlTetrahedra, tet)])
    want_f = O.extract_items(3, [(G.GmlTetrahedra, tet)])
    with G.Gml() as g:
        set_vertices(g, nv)
        set_block(g, G.GmlTetrahedra, tet, np.zeros(nt, np.int32))
        ne, pe = extract_logged(g, "edges", capfd)
        nf, pf = extract_logged(g, "faces", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        f = get_block(g, G.GmlTriangles, nf, 3)
    assert pe == "sorted" and pf == "sorted"
    assert np.array_equal(e[0], want_e[0]) and np.array_equal(e[1], want_e[1])
    assert np.array_equal(f[0], want_f[0]) and np.array_equal(f[1], want_f[1])


def test_gpu_extract_two_word_keys_against_oracle(capfd):
    # the 40^3 cube's node ids spread 64 apart (4.4M vertex slots, 23-bit ids):
    # triangle keys take two words; a work-group's nodes still span a narrow
    # range, so the block pass packs them relative to its lowest node (groups
    # whose range is too wide take the round-based pass); a two-row triangle
    # table already there keeps the call off the vertex-ball path; tables
    # bit-identical to the oracle
    crd, ref, tet = G.cube_mesh(40, 0.2, 21)
    tet = np.ascontiguousarray(tet * 64)
    nv = len(crd) * 64
    of = np.array([[tet[3, 0], tet[3, 2], tet[3, 1]], [tet[8, 1], tet[8, 0], tet[8, 3]]], np.int32)
    oref = np.array([9, 10], np.int32)
    want = O.extract_items(3, [(G.GmlTetrahedra, tet)], of, oref)
    with G.Gml() as g:
        set_vertices(g, nv)
        set_block(g, G.GmlTetrahedra, tet, np.zeros(len(tet), np.int32))
        set_block(g, G.GmlTriangles, of, oref)
        nf, pf = extract_logged(g, "faces", capfd)
        f = get_block(g, G.GmlTriangles, nf, 3)
    assert pf == "bucketed" and nf == len(want[0])
    assert np.array_equal(f[0], want[0]) and np.array_equal(f[1], want[1])


@pytest.mark.parametrize("nt", [100, 250])
def test_gpu_extract_wide_ball_fan(capfd, nt):
    # a fan of nt tets (0, i, i+1, i+2): vertex 0 is the lowest node of
    # nt + 2 edges and 2 nt + 1 faces, more than the vertex-ball walk's key
    # list holds, so its walk splits them into classes; its ball is within the
    # walk's degree bound; tables equal the oracle's
    tet = np.array([[0, i, i + 1, i + 2] for i in range(1, nt + 1)], np.int32)
    nv = nt + 3
    want_e = O.extract_items(2, [(G.GmlTetrahedra, tet)])
    want_f = O.extract_items(3, [(G.GmlTetrahedra, tet)])
    with G.Gml() as g:
        set_vertices(g, nv)
        set_block(g, G.GmlTetrahedra, tet, np.zeros(nt, np.int32))
        ne, pe = extract_logged(g, "edges", capfd)
        nf, pf = extract_logged(g, "faces", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        f = get_block(g, G.GmlTriangles, nf, 3)
    assert pe == "vertex-ball" and pf == "vertex-ball"
    assert np.array_equal(e[0], want_e[0]) and np.array_equal(e[1], want_e[1])
    assert np.array_equal(f[0], want_f[0]) and np.array_equal(f[1], want_f[1])


def test_gpu_extract_degenerate_tet_leaves_ball_path(capfd):
    # a 6^3 cube with one tet that repeats a vertex: the vertex-ball kernel
    # reports it and the general path builds the tables, equal to the oracle's
    crd, ref, tet = G.cube_mesh(6, 0.1, 5)
    tet = tet.copy()
    tet[len(tet) // 2, 3] = tet[len(tet) // 2, 1]
    want_e = O.extract_items(2, [(G.GmlTetrahedra, tet)])
    want_f = O.extract_items(3, [(G.GmlTetrahedra, tet)])
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        ne, pe = extract_logged(g, "edges", capfd)
        nf, pf = extract_logged(g, "faces", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        f = get_block(g, G.GmlTriangles, nf, 3)
    assert pe in ("bucketed", "sorted") and pf in ("bucketed", "sorted")
    assert np.array_equal(e[0], want_e[0]) and np.array_equal(e[1], want_e[1])
    assert np.array_equal(f[0], want_f[0]) and np.array_equal(f[1], want_f[1])


def test_gpu_extract_ball_path_sparse_ids(capfd):
    # the 40^3 cube's node ids spread 64 apart: 63 of every 64 vertices are in
    # no tet (empty balls); tet-only, so the vertex-ball path runs; tables
    # equal the oracle's
    crd, ref, tet = G.cube_mesh(40, 0.2, 23)
    tet = np.ascontiguousarray(tet * 64)
    nv = len(crd) * 64
    want_e = O.extract_items(2, [(G.GmlTetrahedra, tet)])
    want_f = O.extract_items(3, [(G.GmlTetrahedra, tet)])
    with G.Gml() as g:
        set_vertices(g, nv)
        set_block(g, G.GmlTetrahedra, tet, np.zeros(len(tet), np.int32))
        ne, pe = extract_logged(g, "edges", capfd)
        nf, pf = extract_logged(g, "faces", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        f = get_block(g, G.GmlTriangles, nf, 3)
    assert pe == "vertex-ball" and pf == "vertex-ball"
    assert np.array_equal(e[0], want_e[0]) and np.array_equal(e[1], want_e[1])
    assert np.array_equal(f[0], want_f[0]) and np.array_equal(f[1], want_f[1])


@pytest.mark.parametrize("nv,nt", [(100, 2000), (3000, 20000)])
def test_gpu_extract_ball_path_random_soup(capfd, nv, nt):
    # random tets (four distinct vertices each) over few vertices, plus a
    # "book" of 50 tets around the face (0, 1, 2), in shuffled order: faces
    # and edges shared by any number of tets, vertices owning up to hundreds
    # of distinct items (the walk splits them into classes); tables equal the
    # oracle's
    rng = np.random.default_rng(nv)
    soup = [rng.choice(nv, 4, replace=False) for _ in range(nt)]
    book = [[0, 1, 2, 3 + i] for i in range(50)]
    tet = np.array(soup + book, np.int32)[rng.permutation(nt + 50)]
    tet = np.ascontiguousarray(tet)
    nt = len(tet)
    want_e = O.extract_items(2, [(G.GmlTetrahedra, tet)])
    want_f = O.extract_items(3, [(G.GmlTetrahedra, tet)])
    with G.Gml() as g:
        set_vertices(g, nv)
        set_block(g, G.GmlTetrahedra, tet, np.zeros(nt, np.int32))
        ne, pe = extract_logged(g, "edges", capfd)
        nf, pf = extract_logged(g, "faces", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        f = get_block(g, G.GmlTriangles, nf, 3)
    assert pe == "vertex-ball" and pf == "vertex-ball"
    assert np.array_equal(e[0], want_e[0]) and np.array_equal(e[1], want_e[1])
    assert np.array_equal(f[0], want_f[0]) and np.array_equal(f[1], want_f[1])


def _drop(g, typ):
    idx = C.c_int(0)
    if g.L.GmlGetMeshInfo(g.h, typ, None, C.byref(idx)):
        assert g.L.GmlFreeData(g.h, idx.value)


def _rewrite_tets(g, tet):
    # new contents for the existing tet table (same count)
    nt = len(tet)
    tref = np.zeros(nt, np.int32)
    assert g.L.GmlSetDataBlock(g.h, G.GmlTetrahedra, 0, nt - 1, G.ptr(tet), C.c_void_p(tet.ctypes.data + (nt - 1) * 16),
                               G.ptr(tref), C.c_void_p(tref.ctypes.data + (nt - 1) * 4))


def test_gpu_extract_ball_follows_table_changes(capfd):
    # the vertex -> tet ball kept by a first extraction is rebuilt once the
    # tet table has new contents or the vertex count has changed; a ball that
    # a caller holds (GmlxBuildLink) is left as it is and the general path
    # runs; every table equals the oracle's for the current mesh
    crd, ref, tet = G.cube_mesh(12, 0.2, 5)
    rng = np.random.default_rng(8)
    pv = rng.permutation(len(crd))
    inv = np.empty_like(pv)
    inv[pv] = np.arange(len(crd))
    tet2 = np.ascontiguousarray(inv[tet[rng.permutation(len(tet))]].astype(np.int32))
    want1 = O.extract_items(2, [(G.GmlTetrahedra, tet)])
    want2 = O.extract_items(2, [(G.GmlTetrahedra, tet2)])
    wantf = O.extract_items(3, [(G.GmlTetrahedra, tet)])

    def edges(g, want, path):
        _drop(g, G.GmlEdges)
        ne, pe = extract_logged(g, "edges", capfd)
        e = get_block(g, G.GmlEdges, ne, 2)
        assert pe in path
        assert np.array_equal(e[0], want[0]) and np.array_equal(e[1], want[1])

    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        edges(g, want1, ("vertex-ball",))
        _rewrite_tets(g, tet2)                     # same table, new contents
        edges(g, want2, ("vertex-ball",))
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        nf, pf = extract_logged(g, "faces", capfd)
        assert pf == "vertex-ball"
        set_vertices(g, 2 * len(crd))              # a larger vertex table
        _drop(g, G.GmlTriangles)
        nf, pf = extract_logged(g, "faces", capfd)
        f = get_block(g, G.GmlTriangles, nf, 3)
        assert pf == "vertex-ball"
        assert np.array_equal(f[0], wantf[0]) and np.array_equal(f[1], wantf[1])
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        assert g.L.GmlxBuildLink(g.h, G.GmlVertices, G.GmlTetrahedra)   # held by the caller
        edges(g, want1, ("vertex-ball",))
        _rewrite_tets(g, tet2)
        edges(g, want2, ("bucketed", "sorted"))
        # the built-in ball minimum refuses the stale ball it would read
        q = g.L.GmlNewSolutionData(g.h, G.GmlTetrahedra, 1, G.GmlFlt, b"q")
        o = g.L.GmlNewSolutionData(g.h, G.GmlVertices, 1, G.GmlFlt, b"o")
        assert q and o
        assert g.L.GmlxBallMin(g.h, q, o) == -3
