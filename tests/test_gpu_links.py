"""Edge and face shells, downlinks and neighbours built on the GPU (sorted
item keys, gml_items.hip) against the host-only handle's builders (checked
against the reference-hash restatement in test_links.py) and against the
restatement directly: bit-identical degrees, rows and tables."""
import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O
from tests.test_links import (EDG, HEX, QAD, TET, TRI, hex_mesh, link_rows, tet_mesh)

pytestmark = pytest.mark.gpu


def links(host_only, mesh, n, specs):
    """{spec: result} for uplinks (deg, ball), downlinks and neighbours."""
    out = {}
    with G.Gml(host_only=host_only) as g:
        tabs = mesh(g, n)
        for kind, s, d in specs:
            if kind == "up":
                assert g.L.GmlxBuildLink(g.h, s, d)
                ns = len(tabs[s])
                info = g.link_info(s, d)
                width = max(info[1], info[3], 1)
                out[(kind, s, d)] = g.uplink_table(s, d, ns, max(width, 64))
            elif kind == "down":
                assert g.L.GmlxBuildLink(g.h, s, d)
                idx = g.L.GmlxGetLinkData(g.h, s, d)
                nl, w, _, _ = g.link_info(s, d)
                out[(kind, s, d)] = link_rows(g, idx, nl, w)
            else:
                idx = g.L.GmlSetNeighbours(g.h, s)
                assert idx
                nl, w, _, _ = g.link_info(s, s)
                out[(kind, s, s)] = link_rows(g, idx, nl, w)
    return out, tabs


def tet_tabs(g, n):
    tet, edg, tri = tet_mesh(g, n)
    return {TET: tet, EDG: edg, TRI: tri}


def hex_tabs(g, n):
    h, edg, qad = hex_mesh(g, n)
    return {HEX: h, EDG: edg, QAD: qad}


TET_SPECS = [("up", EDG, TET), ("up", TRI, TET), ("down", TET, EDG), ("down", TET, TRI), ("ngb", TET, TET),
             ("ngb", TRI, TRI), ("ngb", EDG, EDG), ("up", EDG, TRI), ("down", TRI, EDG)]
HEX_SPECS = [("up", QAD, HEX), ("up", EDG, HEX), ("down", HEX, QAD), ("down", HEX, EDG), ("ngb", HEX, HEX),
             ("up", EDG, QAD), ("ngb", QAD, QAD)]


@pytest.mark.parametrize("mesh,n,specs", [(tet_tabs, 2, TET_SPECS), (tet_tabs, 5, TET_SPECS),
                                          (hex_tabs, 3, HEX_SPECS)])
def test_gpu_links_same_as_host(mesh, n, specs):
    dev, tabs = links(False, mesh, n, specs)
    host, _ = links(True, mesh, n, specs)
    for k in specs:
        key = k if k[0] != "ngb" else ("ngb", k[1], k[1])
        a, b = dev[key], host[key]
        if k[0] == "up":
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), k
        else:
            assert np.array_equal(a, b), k
    if n <= 3:
        # and against the reference-hash restatement directly
        for kind, s, d in specs:
            if kind == "up":
                want = O.ref_shells(tabs[s], s, tabs[d], d)
                deg, ball = dev[(kind, s, d)]
                assert list(deg) == [len(r) for r in want]
                assert all(list(ball[i, :len(r)]) == r for i, r in enumerate(want))
            elif kind == "down":
                got = dev[(kind, s, d)]
                assert np.array_equal(got, O.ref_downlink(tabs[s], s, tabs[d], d, got.shape[1]))
            else:
                got = dev[("ngb", s, s)]
                assert np.array_equal(got, O.ref_neighbours(tabs[s], s, got.shape[1]))


def test_gpu_links_large_cube_closed_forms():
    import time
    n = 48
    with G.Gml() as g:
        crd, ref, tet = G.cube_mesh(n, 0.1, 4)
        g.set_mesh(crd, ref, tet)
        T = len(tet)
        ne = g.L.GmlExtractEdges(g.h)
        nf = g.L.GmlExtractFaces(g.h)
        t0 = time.perf_counter()
        assert g.L.GmlxBuildLink(g.h, TRI, TET)
        t1 = time.perf_counter()
        assert g.L.GmlxBuildLink(g.h, TET, TRI)
        t2 = time.perf_counter()
        idx = g.L.GmlSetNeighbours(g.h, TET)
        t3 = time.perf_counter()
        print("%d tets: tri->tet shells %.1f ms, tet->tri downlink %.1f ms, tet neighbours %.1f ms"
              % (T, 1e3 * (t1 - t0), 1e3 * (t2 - t1), 1e3 * (t3 - t2)))
        deg, ball = g.uplink_table(TRI, TET, nf, 2)
        assert deg.sum() == 4 * T and (deg == 1).sum() == 12 * n * n
        down = g.download_raw(g.L.GmlxGetLinkData(g.h, TET, TRI), np.int32, (T, 4))
        ngb = g.download_raw(idx, np.int32, (T, 4))
    # every tet's four faces found, each interior face's two tets see each other
    assert (down >= 0).all() and len(np.unique(down)) == nf
    t = np.repeat(np.arange(T), 4)
    k = ngb.ravel()
    inner = k > 0
    back = ngb[k[inner]]
    assert (back == t[inner][:, None]).any(axis=1).all()
    # zero entries: the 12 n^2 boundary faces, plus one per tet across a face from tet 0
    assert (k == 0).sum() == 12 * n * n + (ngb[0] != 0).sum()


def test_gpu_links_reference_sample_mesh(tmp_path):
    # the reference's naca012 sample (72,783 triangles, its boundary edges
    # with refs, then the extracted interior edges): edge -> triangle shells,
    # triangle -> edge downlinks and triangle neighbours, device against the
    # host-only builders
    import lzma
    import os
    from tests.test_meshb import GMF, imp, mesh_tables
    data = lzma.open(os.path.join(os.path.dirname(__file__), "golden", "naca012_2D_triangles.meshb.xz")).read()
    path = str(tmp_path / "naca.meshb")
    open(path, "wb").write(data)
    specs = [("up", EDG, TRI), ("down", TRI, EDG), ("ngb", TRI, TRI)]

    def sample(g, n):
        assert imp(g, path, GMF["Vertices"], GMF["Edges"], GMF["Triangles"]) == 3
        assert g.L.GmlExtractEdges(g.h) == 109407
        t = mesh_tables(g)
        return {EDG: t["Edges"][0], TRI: t["Triangles"][0]}

    dev, _ = links(False, sample, 0, specs)
    host, _ = links(True, sample, 0, specs)
    for k in specs:
        key = k if k[0] != "ngb" else ("ngb", k[1], k[1])
        a, b = dev[key], host[key]
        if k[0] == "up":
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), k
            assert a[0].min() >= 1 and a[0].max() == 2           # a boundary edge has one triangle
        else:
            assert np.array_equal(a, b), k
