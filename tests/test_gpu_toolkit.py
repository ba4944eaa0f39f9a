"""GPU parity of user bodies that call the geometric toolkit (CalTetQal,
CalTriQal, CalTetVol, CalTriSrf; sources/toolkit.cl:79-228) and of the
vertex -> triangle ball pattern of the reference's BallQuality example
(examples/BallQuality/BallQuality.c:95-139: a triangle kernel, then a vertex
kernel reading the triangle values through the vertex ball).  The bodies are
written here; the expected values come from the oracle (orc_caltetqal_f64,
orc_caltriqal_f64, orc_ref_compute_quality_f64, orc_vertex_ball,
orc_ball_min) and from textbook determinants and cross products."""
import ctypes as C
import lzma
import os

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

# per tet: the toolkit quality (with its repeated edge), the reference
# MeshQuality example's quality (max squared edge, the two faces (0,1,2) and
# (0,1,3) only, 7.348469 * volume / (area * max edge)), the volume and the
# sum of the four face areas
TOOLKIT_TET = """{
   float4 a = VerCrd[0], b = VerCrd[1], c = VerCrd[2], d = VerCrd[3];
   float  m = max(max(max(DisPow(a, b), DisPow(a, c)), max(DisPow(a, d), DisPow(b, c))),
                  max(DisPow(b, d), DisPow(c, d)));
   float  v  = CalTetVol(a, b, c, d);
   float  s2 = CalTriSrf(a, b, c) + CalTriSrf(a, b, d);
   float  s4 = s2 + CalTriSrf(a, c, d) + CalTriSrf(b, c, d);
   tk = (float4)(CalTetQal(a, b, c, d), 7.348469f * v / (s2 * sqrt(m)), v, s4);
}"""

# triangle kernel: quality and area of each triangle
TRI_QUALITY = """{
   tq = CalTriQal(VerCrd[0], VerCrd[1], VerCrd[2]);
   ts = CalTriSrf(VerCrd[0], VerCrd[1], VerCrd[2]);
}"""

# vertex kernel over the vertex -> triangle ball: the worst triangle quality
# of the ball and the area-weighted mean quality (tsNul and tqNul are 0 past
# the degree, so the sums run to the padded width)
BALL_QUALITY = """{
   float lo = tq[0], num = 0.f, den = 0.f;
   for(int i = 0; i < tqDegMax; i++)
   {
      if(i < tqDeg)
         lo = min(lo, tq[i]);
      num += ts[i] * tq[i];
      den += ts[i];
   }
   vq = lo;
   vw = den > 0.f ? num / den : 0.f;
}"""


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def test_toolkit_tet_formulas_vs_oracle():
    crd, ref, tet = G.cube_mesh(10, 0.2, 31)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        tk = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "tk")
        k = g.compile(TOOLKIT_TET, "toolkit_tet", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (tk, G.GmlWriteMode, 0)])
        g.launch(k)
        out = g.download_raw(tk, np.float32, (len(tet), 4)).astype(np.float64)
    P = crd[:, :3].astype(np.float64)
    L = O.lib()
    qt = np.array([L.orc_caltetqal_f64(*[O.p(np.ascontiguousarray(P[v])) for v in t]) for t in tet])
    qr = np.array([L.orc_ref_compute_quality_f64(*[O.p(np.ascontiguousarray(P[v])) for v in t]) for t in tet])
    a, b, c, d = (P[tet[:, j]] for j in range(4))
    vol = 0.166666 * np.linalg.det(np.stack([b - a, c - a, d - a], axis=1))
    area = sum(0.5 * np.linalg.norm(np.cross(y - x, z - x), axis=1)
               for x, y, z in ((a, b, c), (a, b, d), (a, c, d), (b, c, d)))
    errs = [rel(out[:, 0], qt).max(), rel(out[:, 1], qr).max(), rel(out[:, 2], vol).max(), rel(out[:, 3], area).max()]
    print("CalTetQal %.2e, reference quality %.2e, CalTetVol %.2e, face areas %.2e (max relative errors)" % tuple(errs))
    assert max(errs) <= 1e-5
    assert qt.min() > 0 and qr.min() > 0          # the jittered cube has no inverted tets


def triangle_handle(g, crd3, tri):
    nv, nt = len(crd3), len(tri)
    L, h = g.L, g.h
    vi = L.GmlNewMeshData(h, G.GmlVertices, nv)
    ti = L.GmlNewMeshData(h, G.GmlTriangles, nt)
    assert vi and ti
    crd = np.zeros((nv, 4), np.float32)
    crd[:, :3] = crd3
    r = np.zeros(max(nv, nt), np.int32)
    tri = np.ascontiguousarray(tri, np.int32)
    assert L.GmlSetDataBlock(h, G.GmlVertices, 0, nv - 1, G.ptr(crd), C.c_void_p(crd.ctypes.data + (nv - 1) * 16),
                             G.ptr(r), C.c_void_p(r.ctypes.data + (nv - 1) * 4))
    assert L.GmlSetDataBlock(h, G.GmlTriangles, 0, nt - 1, G.ptr(tri), C.c_void_p(tri.ctypes.data + (nt - 1) * 12),
                             G.ptr(r), C.c_void_p(r.ctypes.data + (nt - 1) * 4))
    return vi, ti


def ball_quality(g, vi, ti, nv, nt):
    """(triangle quality, triangle area, vertex ball min, vertex weighted
    mean, ball degrees, ball rows) through the generated kernels."""
    tq = g.solution(G.GmlTriangles, 1, G.GmlFlt, "tq")
    ts = g.solution(G.GmlTriangles, 1, G.GmlFlt, "ts")
    vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
    vw = g.solution(G.GmlVertices, 1, G.GmlFlt, "vw")
    kt = g.compile(TRI_QUALITY, "tri_quality", G.GmlTriangles,
                   [(vi, G.GmlReadMode, 0), (tq, G.GmlWriteMode, 0), (ts, G.GmlWriteMode, 0)])
    kv = g.compile(BALL_QUALITY, "ball_quality", G.GmlVertices,
                   [(ti, G.GmlReadMode, 0), (tq, G.GmlReadMode, 0), (ts, G.GmlReadMode, 0),
                    (vq, G.GmlWriteMode, 0), (vw, G.GmlWriteMode, 0)])
    g.launch(kt)
    g.launch(kv)
    deg, ball = g.uplink_table(G.GmlVertices, G.GmlTriangles, nv, 256)
    return (g.download_raw(tq, np.float32, (nt,)), g.download_raw(ts, np.float32, (nt,)),
            g.download_raw(vq, np.float32, (nv,)), g.download_raw(vw, np.float32, (nv,)), deg, ball)


def check_ball_quality(crd3, tri, res):
    tq, ts, vq, vw, deg, ball = res
    nv = len(crd3)
    P = crd3.astype(np.float64)
    L = O.lib()
    q64 = np.array([L.orc_caltriqal_f64(*[O.p(np.ascontiguousarray(P[v])) for v in t]) for t in tri])
    assert rel(tq, q64).max() <= 1e-5
    # the GPU-built ball, bit-exact against the reference's chain order
    d0, b0, md = O.vertex_ball(tri, nv, 256)
    assert np.array_equal(deg, d0) and np.array_equal(ball, b0)
    # ball minimum bit-exact on the GPU's own triangle qualities
    assert np.array_equal(vq, O.ball_min(d0, b0, tq))
    # area-weighted mean against FP64 on the GPU's own values
    num = np.zeros(nv)
    den = np.zeros(nv)
    np.add.at(num, tri.ravel(), np.repeat(ts.astype(np.float64) * tq, 3))
    np.add.at(den, tri.ravel(), np.repeat(ts.astype(np.float64), 3))
    want = np.where(den > 0, num / np.maximum(den, 1e-300), 0.0)
    assert rel(vw, want).max() <= 1e-5
    return md


def test_ball_quality_reference_sample_mesh(tmp_path):
    # the reference's naca012 sample: 36,624 vertices, 72,783 triangles
    from tests.test_meshb import GMF, imp, mesh_tables
    data = lzma.open(os.path.join(os.path.dirname(__file__), "golden", "naca012_2D_triangles.meshb.xz")).read()
    path = str(tmp_path / "naca.meshb")
    open(path, "wb").write(data)
    with G.Gml() as g:
        assert imp(g, path, GMF["Vertices"], GMF["Triangles"]) == 2
        t = mesh_tables(g)
        crd, tri = t["Vertices"][0], t["Triangles"][0]
        vi = C.c_int()
        ti = C.c_int()
        assert g.L.GmlGetMeshInfo(g.h, G.GmlVertices, None, C.byref(vi))
        assert g.L.GmlGetMeshInfo(g.h, G.GmlTriangles, None, C.byref(ti))
        res = ball_quality(g, vi.value, ti.value, len(crd), len(tri))
    md = check_ball_quality(crd[:, :3], tri, res)
    print("naca012: %d triangles, max vertex -> triangle degree %d" % (len(tri), md))


def test_ball_quality_high_degree_soup():
    # a triangle soup whose hub vertices lie in up to 150 triangles: the
    # vertex -> triangle base rows are 8 wide, so the high rows are exercised
    rng = np.random.default_rng(17)
    nv, nt = 4000, 12000
    crd3 = rng.uniform(0.0, 1.0, size=(nv, 3)).astype(np.float32)
    tri = np.empty((nt, 3), np.int32)
    for i in range(nt):
        tri[i] = rng.choice(nv, 3, replace=False)
    hubs = [0, 1, 2, 3, 4]
    for j, hub in enumerate(hubs):
        rows = rng.choice(np.flatnonzero((tri != hub).all(axis=1)), 30 * (j + 1), replace=False)
        tri[rows, rng.integers(3, size=len(rows))] = hub
    tri = tri[(tri[:, 0] != tri[:, 1]) & (tri[:, 1] != tri[:, 2]) & (tri[:, 0] != tri[:, 2])]
    with G.Gml() as g:
        vi, ti = triangle_handle(g, crd3, tri)
        res = ball_quality(g, vi, ti, nv, len(tri))
    md = check_ball_quality(crd3, tri, res)
    print("soup: %d triangles, max degree %d" % (len(tri), md))
    assert md > 100


def test_opencl_vector_literals_on_the_gpu():
    # (T)(a, b, ...) builds a vector from scalar and vector parts in order
    crd, ref, tet = G.cube_mesh(3, 0.2, 2)
    bod = """{
   float2 h = (float2)(VerCrd[0].x, 2.f);
   int4   n = (int4)(TetIdx, (int2)(5, 6), 7);
   tk = (float4)(h, (float2)(3.f, (float)n.x)) + (float4)(0.f, 0.f, 0.f, (float)(n.y + n.z + n.w));
}"""
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        tk = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "tk")
        k = g.compile(bod, "vlit", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (tk, G.GmlWriteMode, 0)])
        g.launch(k)
        out = g.download_raw(tk, np.float32, (len(tet), 4))
    want = np.stack([crd[tet[:, 0], 0], np.full(len(tet), 2.0), np.full(len(tet), 3.0),
                     np.arange(len(tet)) + 18.0], axis=1).astype(np.float32)
    assert np.array_equal(out, want)
