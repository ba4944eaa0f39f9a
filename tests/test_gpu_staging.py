"""LDS-staged uplink gathers of generated kernels (GmlxSetStaging).

The same body compiled with staging (the default) and without must give
bit-identical results: staging changes where the uplink data are read from
(a per-tile LDS copy loaded with coalesced loads), not which values are read
or in which order.  Covered: float4 and float uplink data, a body that also
names the ball entries themselves (they are then read as well), a random
tet soup whose high-degree rows run in the unstaged twin kernel, and
isolated vertices filling whole tiles with nothing to stage."""
import ctypes as C
import os

import numpy as np
import pytest

import gmlib_amd as G

pytestmark = pytest.mark.gpu
KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")

SUM_BODY = """{
   float4 s = (float4)(0.f);
   for(int i = 0; i < BaryDegMax; i++)
      s += Bary[i];
   Out = s;
}"""

IDS_BODY = """{
   float4 s = (float4)(0.f);
   for(int i = 0; i < BaryDeg; i++)
      s += Bary[i] * (float)(i + 1);
   Out = s;
   First = VerTetBal[0].s0 + VerTetBal[0].s1;
}"""


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


def src_of(g, k):
    return g.L.GmlxGetKernelSource(g.h, k).decode()


def both(g, src, name, typ, args):
    """(unstaged, staged) kernels of one body."""
    L = G.lib()
    old = L.GmlxSetStaging(g.h, 0)
    try:
        k0 = g.compile(src, name, typ, args)
        L.GmlxSetStaging(g.h, 1)
        k1 = g.compile(src, name, typ, args)
    finally:
        L.GmlxSetStaging(g.h, old)
    assert "gml_lds" not in src_of(g, k0) and "gml_lds" in src_of(g, k1)
    return k0, k1


def run_pair(g, kernels, out_idx, shapes):
    """Launch each kernel (several times, so the shape tuner finishes) and
    return the downloaded outputs."""
    res = []
    for k in kernels:
        for _ in range(9):
            g.launch(k)
        res.append([g.download_raw(i, dt, sh) for i, dt, sh in zip(out_idx, *zip(*shapes))])
    return res


def mesh_fields(g, crd, ref, tet):
    vi, ti = g.set_mesh(crd, ref, tet)
    g.new_parameters(16, body("smooth_params"))
    bary = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
    qal = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
    ks = g.compile(body("smooth_scatter"), "smooth_scatter", G.GmlTetrahedra,
                   [(vi, G.GmlReadMode, 0), (bary, G.GmlWriteMode, 0)])
    kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra,
                   [(vi, G.GmlReadMode, 0), (qal, G.GmlWriteMode, 0)])
    g.launch(ks)
    g.launch(kq)
    return vi, bary, qal


@pytest.mark.parametrize("n", [7, 24])
def test_staged_gathers_bit_identical(n):
    crd, ref, tet = G.cube_mesh(n, 0.2, 3)
    V = len(crd)
    with G.Gml() as g:
        vi, bary, qal = mesh_fields(g, crd, ref, tet)
        out = g.solution(G.GmlVertices, 1, G.GmlFlt4, "Out")
        vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
        first = g.solution(G.GmlVertices, 1, G.GmlInt, "First")
        a = run_pair(g, both(g, SUM_BODY, "sumbary", G.GmlVertices, [(bary, G.GmlReadMode, 0), (out, G.GmlWriteMode, 0)]),
                     [out], [(np.float32, (V, 4))])
        assert np.array_equal(a[0][0], a[1][0])
        b = run_pair(g, both(g, body("ball_min"), "ball_min", G.GmlVertices,
                             [(qal, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)]), [vq], [(np.float32, (V,))])
        assert np.array_equal(b[0][0], b[1][0])
        ks = both(g, IDS_BODY, "ids", G.GmlVertices,
                  [(bary, G.GmlReadMode, 0), (out, G.GmlWriteMode, 0), (first, G.GmlWriteMode, 0)])
        assert "gml_VerTetBal[k] = v >> 4;" in src_of(g, ks[1])       # the body names the ball: entries read
        c = run_pair(g, ks, [out, first], [(np.float32, (V, 4)), (np.int32, (V,))])
        assert np.array_equal(c[0][0], c[1][0]) and np.array_equal(c[0][1], c[1][1])
        for k in ks:
            grp, rm = C.c_int(), C.c_int()
            assert g.L.GmlxGetKernelShape(g.h, k, C.byref(grp), C.byref(rm))
            assert grp.value in (64, 128, 256)


def test_staged_with_high_degree_twin():
    rng = np.random.default_rng(5)
    nv = 64
    tet = np.array([rng.choice(nv, 4, replace=False) for _ in range(700)], np.int32)
    crd = rng.uniform(size=(nv, 4)).astype(np.float32)
    crd[:, 3] = 0
    ref = np.zeros(nv, np.int32)
    with G.Gml() as g:
        vi, bary, qal = mesh_fields(g, crd, ref, tet)
        assert g.L.GmlxBuildLink(g.h, G.GmlVertices, G.GmlTetrahedra)
        info = g.link_info(G.GmlVertices, G.GmlTetrahedra)
        assert info[2] > 0                                   # high-degree rows exist
        vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
        b = run_pair(g, both(g, body("ball_min"), "ball_min", G.GmlVertices,
                             [(qal, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)]), [vq], [(np.float32, (nv,))])
        assert np.array_equal(b[0][0], b[1][0])
        q = g.download_raw(qal, np.float32, (len(tet),))
    # and both equal the brute-force ball minimum
    want = np.array([q[np.any(tet == v, axis=1)].min() if np.any(tet == v) else 0. for v in range(nv)], np.float32)
    assert np.array_equal(b[1][0], want)


def test_staged_isolated_vertices_and_empty_tiles():
    # vertices in no tet after a small cube: whole tiles of rows without
    # ball entries (no staged entities), a partial last tile, rows past the
    # end reading the last row before the barrier; every vertex gets the sum
    # of its ball in ball order (0 for an isolated one), staged or not
    from tests import oracle_lib as O
    crd, ref, tet = G.cube_mesh(6, 0.2, 13)
    extra = 600
    rng = np.random.default_rng(3)
    iso = np.zeros((extra, 4), np.float32)
    iso[:, :3] = rng.uniform(2., 3., size=(extra, 3))
    crd = np.concatenate([crd, iso])
    ref = np.concatenate([ref, np.zeros(extra, np.int32)])
    V = len(crd)
    with G.Gml() as g:
        vi, bary, qal = mesh_fields(g, crd, ref, tet)
        out = g.solution(G.GmlVertices, 1, G.GmlFlt4, "Out")
        ks = both(g, SUM_BODY, "sumbary", G.GmlVertices, [(bary, G.GmlReadMode, 0), (out, G.GmlWriteMode, 0)])
        a = run_pair(g, ks, [out], [(np.float32, (V, 4))])
        b4 = g.download_raw(bary, np.float32, (len(tet), 4))
    assert np.array_equal(a[0][0], a[1][0])
    deg, ball, _ = O.vertex_ball(tet, V, 32)
    assert (deg[-extra:] == 0).all()
    assert np.array_equal(a[1][0], O.bw_vertices_ball(deg, ball, b4))
    assert (a[1][0][-extra:] == 0).all()
