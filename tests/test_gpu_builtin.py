"""The hand-written mesh kernels against the generated drop-in kernels of the
same bodies (bit-exact) and against the oracle."""
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


@pytest.mark.parametrize("n,grp,remap,fused", [(20, 256, 1, 1), (13, 64, 0, 1), (11, 128, 1, 1),
                                                (20, 256, 1, 0), (5, 1024, 1, 0),
                                                (20, 256, 1, 2), (13, 64, 0, 2), (20, 512, 1, 2),
                                                (11, 1024, 0, 2)])
def test_builtin_quality_ballmin_smoothing_match_generated(n, grp, remap, fused):
    tile = grp if fused == 2 else min(grp, 256)        # the operator form takes tiles up to 1024
    crd, ref, tet = G.cube_mesh(n, 0.2, 3)
    omega = 0.2
    res = {}
    for path in ("generated", "builtin"):
        with G.Gml() as g:
            vi, ti = g.set_mesh(crd, ref, tet)
            par = g.new_parameters(16, body("smooth_params"))
            C.cast(par, C.POINTER(C.c_float))[0] = omega
            g.L.GmlUploadParameters(g.h)
            q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
            vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
            b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
            if path == "generated":
                kq = g.compile(body("tet_meanratio"), "q", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
                kb = g.compile(body("ball_min"), "bm", G.GmlVertices, [(q, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)])
                ks = g.compile(body("smooth_scatter"), "s", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (b, G.GmlWriteMode, 0)])
                kg = g.compile(body("smooth_gather"), "gt", G.GmlVertices,
                               [(b, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
                g.launch(kq); g.launch(kb)
                q0 = g.download(q, np.float32, (len(tet),)); v0 = g.download(vq, np.float32, (len(crd),))
                for _ in range(5):
                    g.launch(ks); g.launch(kg)
            else:
                # launch shapes and sweep form are this handle's own
                for w in range(4):
                    assert g.L.GmlxSetBuiltinShape(g.h, w, grp, remap)
                assert g.L.GmlxSetBuiltinShape(g.h, G.GmlxBuiltinSweep, tile, remap)
                assert g.L.GmlxSetSmoothMode(g.h, fused)
                assert g.L.GmlxTetQuality(g.h, q) == 1
                assert g.L.GmlxBallMin(g.h, q, vq) == 1
                q0 = g.download(q, np.float32, (len(tet),)); v0 = g.download(vq, np.float32, (len(crd),))
                assert g.L.GmlxSmoothSweeps(g.h, b, 5, omega) == 1
                nl = C.c_longlong()
                which = G.GmlxBuiltinSweep if fused else G.GmlxBuiltinGather
                assert g.L.GmlxGetBuiltinRunTime(g.h, which, C.byref(nl)) > 0 and nl.value == 5
                if fused:
                    info = (C.c_longlong * 9)()
                    assert g.L.GmlxGetSweepPlanInfo(g.h, info) and info[0] == tile
                    print("plan: tiles %d, tile tets per tet %.3f, tile vertices per vertex %.3f"
                          % (info[1], info[3] / len(tet), info[2] / len(crd)))
            c, r = g.vertices(len(crd))
            res[path] = (q0, v0, c)
    for a, b_ in zip(res["generated"][:2], res["builtin"][:2]):
        assert np.array_equal(a, b_)
    cg, cb = res["generated"][2], res["builtin"][2]
    if fused == 2:
        # operator form: same sweep up to FP32 summation order
        print("operator form vs two kernels: max |dx| %.3e" % np.abs(cg - cb).max())
        assert np.abs(cg - cb).max() <= 2e-6
    else:
        assert np.array_equal(cg, cb)
    q64 = O.tet_quality64(tet, crd)
    assert np.max(np.abs(res["builtin"][0] - q64) / np.maximum(np.abs(q64), 1e-3)) <= 1e-5
    d0, b0, _ = O.vertex_ball(tet, len(crd), 32)
    assert np.array_equal(res["builtin"][1], O.ball_min(d0, b0, res["builtin"][0]))
    c64 = O.smooth64(tet, ref, d0, b0, crd, 5, omega)
    assert np.abs(res["builtin"][2][:, :3] - c64).max() <= 1e-5



@pytest.mark.parametrize("degenerate", [False, True])
def test_operator_sweep_degree_from_counts(degenerate):
    # the operator form derives deg v from its 1-ring counts (each ball tet
    # gives its three other vertex positions); a tet listing a vertex twice
    # adds a self slot, so the sweep stays the mean ball barycentre of the
    # reference's ball (the tet counted once per position)
    crd, ref, tet = G.cube_mesh(6, 0.2, 7)
    if degenerate:
        inner = np.nonzero((ref[tet] == 0).all(axis=1))[0][0]
        a, b, c, _ = tet[inner]
        tet = np.vstack([tet, [[a, a, b, c]]]).astype(np.int32)
    omega = 0.3
    with G.Gml() as g:
        assert g.L.GmlxSetSmoothMode(g.h, 2)
        vi, ti = g.set_mesh(crd, ref, tet)
        bary = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        assert g.L.GmlxSmoothSweeps(g.h, bary, 2, omega) == 1
        info = (C.c_longlong * 9)()
        assert g.L.GmlxGetSweepPlanInfo(g.h, info) and info[8] == 2
        c, _ = g.vertices(len(crd))
    deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
    if degenerate:
        assert deg[a] == np.count_nonzero(tet == a)            # the repeated vertex's ball holds the tet twice
    c64 = O.smooth64(tet, ref, deg, ball, crd, 2, omega)
    assert np.abs(c[:, :3] - c64).max() <= 2e-6


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_builtin_sweeps_follow_a_new_mesh(mode):
    # GmlxSmoothSweeps and GmlxBallMin on one handle, then again after a
    # larger mesh is loaded into the same handle: the vertex ball, the sweep
    # plans and the alternate coordinate buffer follow the new mesh;
    # coordinates within 1e-5 of the oracle's FP32 sweeps, ball minimum
    # bit-exact on the GPU's own qualities
    omega = 0.2
    with G.Gml() as g:
        assert g.L.GmlxSetSmoothMode(g.h, mode)
        for n in (6, 11):
            crd, ref, tet = G.cube_mesh(n, 0.2, 4 + n)
            V, T = len(crd), len(tet)
            g.set_mesh(crd, ref, tet)
            b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
            q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
            vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
            assert g.L.GmlxSmoothSweeps(g.h, b, 3, omega) == 1
            c, _ = g.vertices(V)
            deg, ball, _ = O.vertex_ball(tet, V, 32)
            want = O.smooth32(tet, ref, deg, ball, crd, 3, omega)[:, :3]
            assert float(np.abs(c[:, :3] - want).max()) <= 1e-5
            assert g.L.GmlxTetQuality(g.h, q) == 1
            assert g.L.GmlxBallMin(g.h, q, vq) == 1
            qv = g.download(q, np.float32, (T,))
            assert np.array_equal(g.download(vq, np.float32, (V,)), O.ball_min(deg, ball, qv))
