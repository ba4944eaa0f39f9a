"""Paired form of generated kernels (two rows per thread, reads of both
issued before either body): bit-identical to the one-row form for every
launch shape, including the last work-group's missing second rows and
bodies that return early; the tuner settles on one of the two forms."""
import ctypes as C
import os

import numpy as np
import pytest

import gmlib_amd as G
from gmlib_amd import heat as H

pytestmark = pytest.mark.gpu
KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


SHAPES = [(256, 1), (128, 0), (64, 1)]


def run_forms(build, launch, read):
    """results of one kernel for every (shape, rows) on fresh handles"""
    out = {}
    for grp, remap in SHAPES:
        for rows in (1, 2):
            with G.Gml() as g:
                k, ctx = build(g)
                assert g.L.GmlxSetKernelShape(g.h, k, grp, remap, rows) == 1, (grp, remap, rows)
                launch(g, k, ctx)
                out[(grp, remap, rows)] = read(g, ctx)
    return out


@pytest.mark.parametrize("n", [7, 12])
def test_paired_quality_and_scatter_bit_identical(n):
    crd, ref, tet = G.cube_mesh(n, 0.2, 5)
    T = len(tet)
    for kname, typ, item, out in (("tet_meanratio", G.GmlFlt, (T,), "qal"), ("smooth_scatter", G.GmlFlt4, (T, 4), "Bary")):
        def build(g):
            vi, ti = g.set_mesh(crd, ref, tet)
            q = g.solution(G.GmlTetrahedra, 1, typ, out)
            k = g.compile(body(kname), kname, G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
            assert g.L.GmlxGetKernelPairSource(g.h, k) is not None
            return k, q

        res = run_forms(build, lambda g, k, q: g.launch(k),
                        lambda g, q: g.download_raw(q, np.float32, item))
        ref0 = res[(256, 1, 1)]
        assert np.isfinite(ref0).all()
        for key, v in res.items():
            assert np.array_equal(v, ref0), (kname, key)


def test_paired_early_return_and_partial_group():
    crd, ref, tet = G.cube_mesh(5, 0.2, 2)      # 750 tets: the last group of 512 rows is partial
    T = len(tet)

    def build(g):
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "out")
        k = g.compile("if(TetIdx % 3 == 1) return;\nout = (float)TetIdx + VerCrd[2].x;", "r", G.GmlTetrahedra,
                      [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        return k, q

    res = run_forms(build, lambda g, k, q: g.launch(k), lambda g, q: g.download_raw(q, np.float32, (T,)))
    t = np.arange(T)
    want = np.where(t % 3 == 1, 0.0, t + crd[tet[:, 2], 0]).astype(np.float32)
    for key, v in res.items():
        assert np.array_equal(v, want), key


def test_paired_heat_kernels_and_tuner():
    # face kernels gather their two cells through the tri -> tet shell, cell
    # kernels their faces through the tet -> tri downlink
    crd, vref, tet, tri, side = H.heat_cube(6, 0.2, 4)
    outs = []
    for rows in (1, 2):
        s = H.HeatSolver(crd, vref, tet, tri, side)
        try:
            for k in (s.k_face, s.k_grad, s.k_fgrad, s.k_bal, s.k_step):
                assert s.g.L.GmlxGetKernelPairSource(s.g.h, k) is not None
                assert s.g.L.GmlxSetKernelShape(s.g.h, k, 256, 1, rows) == 1
            s.iterate(15)
            outs.append(s.temperature())
        finally:
            s.close()
    assert np.array_equal(outs[0], outs[1])
    # the tuner tries every form (one row, paired, element tiles: 15 shapes
    # and the first again) and keeps one
    crd, ref, tet = G.cube_mesh(10, 0.2, 1)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile(body("tet_meanratio"), "q", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        for _ in range(15):
            g.launch(k)
        assert g.L.GmlxGetKernelForm(g.h, k) == 0
        g.launch(k)
        assert g.L.GmlxGetKernelForm(g.h, k) in (1, 2, 3)
        assert g.L.GmlxGetKernelRows(g.h, k) in (1, 2, 4, 8)
