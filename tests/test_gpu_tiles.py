"""The element-tile form of generated element kernels and of the hand-written
quality kernel: a work-group takes a tile of consecutive elements, loads the
tile's distinct vertices once into LDS and every element reads its corners
there (GmlxGetKernelForm 3).  Every form of a kernel must give the same
values as its one-row form, the tile plan must follow changes of the element
table, and a table whose tiles reach too many vertices must fall back to the
other forms."""
import ctypes as C
import os

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


def run_forms(g, k, out, shape_out, nlaunch=1):
    """{form: output} for the one-row (1), paired (2) and tile (-1) forms the
    kernel has."""
    L, h = g.L, g.h
    res = {}
    for rows in (1, 2, -1):
        for grp in (256, 128, 64):
            if not L.GmlxSetKernelShape(h, k, grp, 1, rows):
                continue
            for _ in range(nlaunch):
                g.launch(k)
            res[(rows, grp)] = (L.GmlxGetKernelForm(h, k), g.download_raw(out, *shape_out))
    return res


def test_tile_form_generated_quality_and_scatter_match_other_forms():
    crd, ref, tet = G.cube_mesh(24, 0.2, 5)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        ks = g.compile(body("smooth_scatter"), "smooth_scatter", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (b, G.GmlWriteMode, 0)])
        assert g.L.GmlxGetKernelTileSource(g.h, kq) and g.L.GmlxGetKernelTileSource(g.h, ks)
        rq = run_forms(g, kq, q, (np.float32, (len(tet),)))
        rs = run_forms(g, ks, b, (np.float32, (len(tet), 4)))
    for r in (rq, rs):
        forms = {v[0] for v in r.values()}
        assert forms == {1, 2, 3}, forms
        ref_out = r[(1, 256)][1]
        for key, (form, out) in r.items():
            assert np.array_equal(out, ref_out), (key, form)
    q64 = O.tet_quality64(tet, crd)
    assert np.max(np.abs(rq[(-1, 128)][1] - q64) / np.maximum(np.abs(q64), 1e-3)) <= 1e-5
    bary = crd[tet].mean(axis=1)
    assert np.abs(rs[(-1, 128)][1] - bary).max() <= 1e-6


def test_tile_form_tuned_and_builtin_quality():
    crd, ref, tet = G.cube_mesh(30, 0.2, 6)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        q2 = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "q2")
        kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        for _ in range(25):                              # every shape of the tuner, then steady state
            g.launch(kq)
        assert g.L.GmlxGetKernelForm(g.h, kq) in (1, 2, 3)
        qg = g.download_raw(q, np.float32, (len(tet),))
        out = {}
        for per in (0, 1, 2, 4):
            assert g.L.GmlxSetBuiltinItems(g.h, G.GmlxBuiltinQuality, per)
            for grp in (64, 128, 256):
                assert g.L.GmlxSetBuiltinShape(g.h, G.GmlxBuiltinQuality, grp, 1)
                assert g.L.GmlxTetQuality(g.h, q2) == 1
                out[(per, grp)] = g.download_raw(q2, np.float32, (len(tet),))
    for key, o in out.items():
        assert np.array_equal(o, qg), key


def test_tile_plan_follows_table_changes():
    # the tile plan of the tet table is rebuilt when the table changes: the
    # tets reversed (the same mesh in another order) give the reversed result
    crd, ref, tet = G.cube_mesh(12, 0.2, 9)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        q2 = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "q2")
        kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        assert g.L.GmlxSetKernelShape(g.h, kq, 128, 1, -1)
        g.launch(kq)
        for _ in range(2):                              # the second call on the table uses its plan
            assert g.L.GmlxTetQuality(g.h, q2) == 1
        a = g.download_raw(q, np.float32, (len(tet),))
        a2 = g.download_raw(q2, np.float32, (len(tet),))
        rev = np.ascontiguousarray(tet[::-1])
        tref = np.zeros(len(tet), np.int32)
        assert g.L.GmlSetDataBlock(g.h, G.GmlTetrahedra, 0, len(tet) - 1, G.ptr(rev),
                                   C.c_void_p(rev.ctypes.data + (len(tet) - 1) * 16), G.ptr(tref),
                                   C.c_void_p(tref.ctypes.data + (len(tet) - 1) * 4))
        g.launch(kq)
        for _ in range(2):
            assert g.L.GmlxTetQuality(g.h, q2) == 1
        b = g.download_raw(q, np.float32, (len(tet),))
        b2 = g.download_raw(q2, np.float32, (len(tet),))
        assert g.L.GmlxGetKernelForm(g.h, kq) == 3
    assert np.array_equal(b, a[::-1]) and np.array_equal(b2, a2[::-1]) and np.array_equal(a, a2)


def hex_grid(n):
    """n^3 unit-cube hexes (bottom face 0-3 counter-clockwise, top 4-7)."""
    g = np.linspace(0.0, 1.0, n + 1)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    crd = np.zeros(((n + 1) ** 3, 4), np.float32)
    crd[:, 0], crd[:, 1], crd[:, 2] = x.ravel(), y.ravel(), z.ravel()
    v = np.arange((n + 1) ** 3).reshape(n + 1, n + 1, n + 1)
    c = v[:-1, :-1, :-1].ravel()
    dx, dy, dz = (n + 1) ** 2, n + 1, 1
    hexes = np.stack([c, c + dx, c + dx + dy, c + dy, c + dz, c + dx + dz, c + dx + dy + dz, c + dy + dz], axis=1)
    return crd, hexes.astype(np.int32)


def test_tile_form_falls_back_on_a_scattered_numbering():
    # hexes in random order: even a tile of 64 consecutive hexes (512 slots)
    # reaches more than 256 vertices, so the table has no tile plan; the
    # tuner keeps the other forms, a pinned tile form runs one row per
    # thread, and the values stay those of the one-row form
    crd2, hexes = hex_grid(12)
    hex2 = np.ascontiguousarray(hexes[np.random.default_rng(0).permutation(len(hexes))])
    body_hex = "{ hv = CalHexVol(VerCrd[0], VerCrd[1], VerCrd[2], VerCrd[3], VerCrd[4], VerCrd[5], VerCrd[6], VerCrd[7]); }"
    with G.Gml() as g:
        L, h = g.L, g.h
        vi = L.GmlNewMeshData(h, G.GmlVertices, len(crd2))
        hi = L.GmlNewMeshData(h, G.GmlHexahedra, len(hex2))
        r = np.zeros(max(len(crd2), len(hex2)), np.int32)
        assert L.GmlSetDataBlock(h, G.GmlVertices, 0, len(crd2) - 1, G.ptr(crd2),
                                 C.c_void_p(crd2.ctypes.data + (len(crd2) - 1) * 16), G.ptr(r), G.ptr(r))
        assert L.GmlSetDataBlock(h, G.GmlHexahedra, 0, len(hex2) - 1, G.ptr(hex2),
                                 C.c_void_p(hex2.ctypes.data + (len(hex2) - 1) * 32), G.ptr(r), G.ptr(r))
        hv = g.solution(G.GmlHexahedra, 1, G.GmlFlt, "hv")
        k = g.compile(body_hex, "hex_volume", G.GmlHexahedra, [(vi, G.GmlReadMode, 0), (hv, G.GmlWriteMode, 0)])
        assert g.L.GmlxGetKernelTileSource(h, k)
        for _ in range(25):
            g.launch(k)
        assert L.GmlxGetKernelForm(h, k) in (1, 2)
        a = g.download_raw(hv, np.float32, (len(hex2),))
        assert L.GmlxSetKernelShape(h, k, 128, 1, -1)
        g.launch(k)
        assert L.GmlxGetKernelForm(h, k) == 1
        b = g.download_raw(hv, np.float32, (len(hex2),))
    assert np.array_equal(a, b)
    # CalHexVol of a unit-grid hex: the toolkit's 8-tet sum (x 0.125)
    assert np.allclose(a, np.float32(1.0 / 12 ** 3), rtol=1e-5)


def test_tile_form_scattered_tets_use_small_tiles():
    # tets in random order with vertices numbered at random still fit
    # 64-tet tiles (at most 256 slots): the tile form stays available and
    # equal to the one-row form
    crd, ref, tet = G.cube_mesh(16, 0.2, 3)
    perm = np.random.default_rng(0).permutation(len(crd))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(crd))
    crd2 = np.ascontiguousarray(crd[perm])
    ref2 = np.ascontiguousarray(ref[perm])
    tet2 = np.ascontiguousarray(inv[tet].astype(np.int32)[np.random.default_rng(1).permutation(len(tet))])
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd2, ref2, tet2)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        res = run_forms(g, kq, q, (np.float32, (len(tet2),)))
        assert g.L.GmlxTetQuality(g.h, q) == 1
        c = g.download_raw(q, np.float32, (len(tet2),))
    assert 3 in {v[0] for v in res.values()}
    for key, (form, out) in res.items():
        assert np.array_equal(out, c), (key, form)
    q64 = O.tet_quality64(tet2, crd2)
    assert np.max(np.abs(c - q64) / np.maximum(np.abs(q64), 1e-3)) <= 1e-5


def test_tile_form_triangles_with_smaller_tiles():
    # a triangle table (int4 rows, the fourth slot 0) of a Hilbert-ordered
    # surface: 512-triangle tiles reach more than 256 vertices, smaller
    # tiles are used; the values equal the one-row form's
    crd, ref, tet = G.cube_mesh(20, 0.2, 8)
    faces = np.sort(np.concatenate([tet[:, [0, 1, 2]], tet[:, [0, 1, 3]], tet[:, [0, 2, 3]], tet[:, [1, 2, 3]]]), axis=1)
    u, cnt = np.unique(faces, axis=0, return_counts=True)
    tri = np.ascontiguousarray(u[cnt == 1].astype(np.int32))       # boundary triangles
    body_tri = "{ ta = CalTriSrf(VerCrd[0], VerCrd[1], VerCrd[2]); }"
    with G.Gml() as g:
        L, h = g.L, g.h
        vi = L.GmlNewMeshData(h, G.GmlVertices, len(crd))
        ti = L.GmlNewMeshData(h, G.GmlTriangles, len(tri))
        r = np.zeros(max(len(crd), len(tri)), np.int32)
        assert L.GmlSetDataBlock(h, G.GmlVertices, 0, len(crd) - 1, G.ptr(crd),
                                 C.c_void_p(crd.ctypes.data + (len(crd) - 1) * 16), G.ptr(r), G.ptr(r))
        assert L.GmlSetDataBlock(h, G.GmlTriangles, 0, len(tri) - 1, G.ptr(tri),
                                 C.c_void_p(tri.ctypes.data + (len(tri) - 1) * 12), G.ptr(r), G.ptr(r))
        ta = g.solution(G.GmlTriangles, 1, G.GmlFlt, "ta")
        k = g.compile(body_tri, "tri_area", G.GmlTriangles, [(vi, G.GmlReadMode, 0), (ta, G.GmlWriteMode, 0)])
        res = run_forms(g, k, ta, (np.float32, (len(tri),)))
    forms = {v[0] for v in res.values()}
    assert 3 in forms, forms
    ref_out = res[(1, 256)][1]
    for key, (form, out) in res.items():
        assert np.array_equal(out, ref_out), (key, form)
    P = crd[:, :3].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(P[tri[:, 1]] - P[tri[:, 0]], P[tri[:, 2]] - P[tri[:, 0]]), axis=1)
    assert np.max(np.abs(ref_out - area) / area) <= 1e-5
