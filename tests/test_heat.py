"""Cell-centered heat solver: the numpy FP64 restatement of the reference
scheme (tests/oracle_lib.py, HeatCellCentered example) pinned by closed
forms, and the engine's generated kernels for it (shells, downlinks, refs,
parameters) on a host-only handle."""
import numpy as np

import gmlib_amd as G
from gmlib_amd import heat as H
from tests import oracle_lib as O


def test_green_gauss_closed_forms():
    crd, vref, tet, tri, side = H.heat_cube(3, 0.25, 5)
    nrm, vol = O.heat_geometry(crd, tet)
    # closed surfaces: the area vectors of every tet sum to zero
    assert np.abs(nrm.sum(axis=1)).max() < 1e-12
    # a linear field at the face centroids: the gradient is exact (up to the 0.166666 volume factor)
    a = np.array([0.3, -1.7, 2.5])
    c = crd[:, :3].astype(np.float64)
    uf = np.stack([(c[tet[:, list(f)]].mean(axis=1) @ a) for f in O.HEAT_FACES], 1) + 4.0
    g = 0.5 * np.einsum("tj,tjk->tk", uf, nrm) / vol[:, None]
    assert np.abs(g - a / (6 * O.CAL_TET_VOL)).max() < 1e-10
    assert (vol > 0).all() and len(tri) == 12 * 9 and set(side) == {1, 2, 3, 4, 5, 6}


def test_links_and_steady_state_of_restatement():
    crd, vref, tet, tri, side = H.heat_cube(2, 0.2, 1)
    full, fref = O.extract_items(3, [(G.GmlTetrahedra, tet)], tri, side)
    shell, down = O.heat_links(len(crd), tet, full)
    assert (shell[:, 0] >= 0).all() and ((shell[:, 1] < 0) == (fref > 0)).all()
    assert (shell[:, 0] > shell[:, 1]).all()                      # descending, as the shells
    u, res, dt = O.heat_iterations(crd, tet, full, fref, H.DEFAULT_BC, 3000)
    assert res[-1] / res[0] < 1e-8
    # sides 4 (z = 0) and 5 (z = 1) at 300 and 250, the others insulated: T = 300 - 50 z, up to
    # the scheme's error on a coarse jittered mesh (face values as cell means)
    zc = crd[tet, 2].astype(np.float64).mean(axis=1)
    slope, icpt = np.linalg.lstsq(np.stack([zc, np.ones_like(zc)], 1), u, rcond=None)[0]
    assert abs(slope + 50) < 1.5 and abs(icpt - 300) < 1 and np.sqrt(((u - 300 + 50 * zc) ** 2).mean()) < 2


def test_heat_kernels_generate_links_and_refs():
    crd, vref, tet, tri, side = H.heat_cube(2, 0.2, 1)
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, vref, tet)
        g.new_parameters(4 * 19, H.body("heat_params"))
        import ctypes as C
        nt = len(tri)
        assert g.L.GmlNewMeshData(g.h, G.GmlTriangles, nt)
        assert g.L.GmlSetDataBlock(g.h, G.GmlTriangles, 0, nt - 1, G.ptr(tri),
                                   C.c_void_p(tri.ctypes.data + (nt - 1) * 12), G.ptr(side),
                                   C.c_void_p(side.ctypes.data + (nt - 1) * 4))
        nf = g.L.GmlExtractFaces(g.h)
        idx = C.c_int(0)
        g.L.GmlGetMeshInfo(g.h, G.GmlTriangles, None, C.byref(idx))
        sol = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "SolTet")
        ext = g.solution(G.GmlTriangles, 1, G.GmlFlt, "SolExt")
        grd = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "GrdTet")
        k1 = g.compile(H.body("heat_face_value"), "heat_face_value", G.GmlTriangles,
                       [(idx.value, G.GmlReadMode | G.GmlRefFlag, 0), (sol, G.GmlReadMode, 0),
                        (ext, G.GmlWriteMode, 0)])
        k2 = g.compile(H.body("heat_cell_gradient"), "heat_cell_gradient", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (ext, G.GmlReadMode, 0), (grd, G.GmlWriteMode, 0)])
        s1 = g.L.GmlxGetKernelSource(g.h, k1).decode()
        s2 = g.L.GmlxGetKernelSource(g.h, k2).decode()
        assert g.link_info(G.GmlTriangles, G.GmlTetrahedra)[1] == 2
        assert g.link_info(G.GmlTetrahedra, G.GmlTriangles)[1] == 4
    assert nf == (4 * len(tet) + 12 * 4) // 2
    assert "TriRef" in s1 and "SolTet[1]" in s1 and "TriTetDeg" in s1
    assert "SolExt[3]" in s2 and "TetTri" in s2
