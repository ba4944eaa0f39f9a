"""GmlCompileKernel on host-only handles: argument classification and the
generated-kernel contract (names, locals, counters, reads, writes of
sources/gmlib3.c:2502-2938), hipRTC compilation for gfx950, host-built links
against the oracle, the high-degree twin kernel, and the reference's own
example bodies compiled unmodified (when the reference tree is present)."""
import ctypes as C
import os

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")
REF = "/root/reference/examples"


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


def src_of(g, k):
    return g.L.GmlxGetKernelSource(g.h, k).decode()


@pytest.fixture
def mesh():
    return G.cube_mesh(4, 0.2, 3, True)


def test_downlink_kernel_contract(mesh):
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra,
                      [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        s = src_of(g, k)
        hdr = s[s.index("// KERNEL HEADER"):]
        # link argument first, then data in user order
        assert hdr.index("int4 *TetVerTab") < hdr.index("float4 *VerCrdTab") < hdr.index("float *qalTab")
        assert "int       TetIdx = cnt + count.s1;" in hdr
        assert "TetVer = __builtin_nontemporal_load(&TetVerTab[ cnt + count.s1 ]);" in hdr
        for i in range(4):
            assert "VerCrd[%d] = VerCrdTab[ TetVer.s%d ];" % (i, i) in hdr
        assert "__builtin_nontemporal_store(qal, &qalTab[ cnt + count.s1 ]);" in hdr
        assert "float4 VerCrd[4];" in hdr


@pytest.mark.parametrize("staged", [0, 1])
def test_uplink_kernel_contract_with_ref_and_params(mesh, staged):
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        G.lib().GmlxSetStaging(g.h, staged)
        vi, ti = g.set_mesh(crd, ref, tet)
        g.new_parameters(16, body("smooth_params"))
        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        k = g.compile(body("smooth_gather"), "smooth_gather", G.GmlVertices,
                      [(b, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
        s = src_of(g, k)
        linkinfo = g.link_info(G.GmlVertices, G.GmlTetrahedra)
    assert "#define   BaryDegMax 32" in s
    assert "int16 VerTetBal[2];" in s and "int VerTetDeg;" in s and "int BaryDeg;" in s
    if staged:
        # the row's own reads, then the tile's loads into LDS, the barrier and
        # the bounds return; the body reads Bary through an LDS accessor by
        # tile-local index; it does not name the ball, so its entries are not read
        h = s[s.index("// KERNEL HEADER"):]
        assert h.index("VerRef = __builtin_nontemporal_load(&VerRefTab[ cnt + count.s1 ]);") \
            < h.index("gml_v_Bary[gml_u] = BaryTab[gml_x[gml_u]];") < h.index("gml_s_Bary[gml_j] = gml_v_Bary[gml_u];") \
            < h.index("__syncthreads();") < h.index("if(cnt >= count.s0)")
        assert "gml_stg<float4, 32> Bary{gml_s_Bary, gml_l_VerTetBal, BaryDeg, BaryNul};" in s
        assert "gml_VerTetBal[k] = v >> 4;" not in s
    else:
        assert "Bary[23] = (BaryDeg >= 24) ? BaryTab[ gml_VerTetBal[23] ] : BaryNul;" in s
        assert "gml_lds" not in s
    for s in [s]:
        assert "VerRef = __builtin_nontemporal_load(&VerRefTab[ cnt + count.s1 ]);" in s
        assert "__builtin_nontemporal_store(VerCrd, &VerCrdTab[ cnt + count.s1 ]);" in s
        assert "__builtin_nontemporal_store(VerRef" not in s
        assert "float omega;" in s
    assert linkinfo == (len(crd), 32, 0, 0)


def test_host_uplink_matches_oracle(mesh):
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        g.set_mesh(crd, ref, tet)
        assert g.L.GmlxBuildLink(g.h, G.GmlVertices, G.GmlTetrahedra)
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, len(crd), 32)
    d0, b0, _ = O.vertex_ball(tet, len(crd), 32)
    assert np.array_equal(deg, d0) and np.array_equal(ball, b0)


def test_high_degree_twin_kernel():
    # a fan of 40 tets around vertex 0 exceeds the 32-wide base rows
    n = 40
    crd = np.zeros((n + 3, 4), np.float32)
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False)
    crd[1:n + 1, 0], crd[1:n + 1, 1] = np.cos(ang), np.sin(ang)
    crd[n + 1, 2], crd[n + 2, 2] = 1.0, -1.0
    tet = []
    for i in range(n):
        a, b = 1 + i, 1 + (i + 1) % n
        tet.append([0, a, b, n + 1] if i % 2 == 0 else [0, b, a, n + 2])
    tet = np.array(tet, np.int32)
    ref = np.zeros(len(crd), np.int32)
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
        k = g.compile(body("ball_min"), "ball_min", G.GmlVertices, [(q, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)])
        # vertex 0 is the first with degree > 32: every row from it on is high
        assert g.link_info(G.GmlVertices, G.GmlTetrahedra) == (0, 32, len(crd), 64)
        assert "#define   qalDegMax 32" in src_of(g, k)
        assert "#define   qalDegMax 64" in src_of(g, k + 1)
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, len(crd), 64)
    d0, b0, md = O.vertex_ball(tet, len(crd), 64)
    assert md == n and np.array_equal(deg, d0) and np.array_equal(ball, b0)


def test_single_precision_literals(mesh):
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        bod = "// 1.5 in a comment\nqal = 0.5 * VerCrd[0].x + 1e-3 + 2. + .25 + 0x1p3 + 3.0f + (float)(1./3.); /* 7.5 */"
        s = src_of(g, g.compile(bod, "lit", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)]))
        user = s[s.index("// USER'S KERNEL CODE"):]
        assert "0.5f * VerCrd[0].x + 1e-3f + 2.f + .25f + 0x1p3 + 3.0f + (float)(1.f/3.f)" in user
        assert "// 1.5 in a comment" in user and "/* 7.5 */" in user


def test_compile_error_is_reported(mesh, capfd):
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        tri = (C.c_int * 6)(vi, G.GmlReadMode, 0, q, G.GmlWriteMode, 0)
        k = g.L.GmlxCompileKernelv(g.h, b"qal = undefined_thing;", b"bad", G.GmlTetrahedra, 2, tri)
        assert k == 0
    assert "undefined_thing" in capfd.readouterr().out


def test_lazy_uplink_accessor_and_voyeurs(mesh):
    # 4 x float4 per tet read through the ball with voyeurs, the item chosen
    # by each entry's local index: only that item is fetched
    from tests.test_gpu_parity import FACE_PULL_GATHER, FACE_PULL_SCATTER
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        o = g.solution(G.GmlTetrahedra, 4, G.GmlFlt4, "Pull")
        r = g.solution(G.GmlVertices, 1, G.GmlFlt, "move")
        ks = g.compile(FACE_PULL_SCATTER, "sca", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (o, G.GmlWriteMode, 0)])
        k = g.compile(FACE_PULL_GATHER, "gat", G.GmlVertices,
                      [(r, G.GmlWriteMode, 0), (o, G.GmlReadMode | G.GmlVoyeurs, 0), (vi, G.GmlReadMode | G.GmlWriteMode, 0)])
        s = src_of(g, k)
        assert ks > 0
        assert "char      VerTetVoy[32];" in s
        assert "VerTetVoy[k] = (char)(v & 7);" in s
        assert "gml_upn<float4, 4, 32> Pull{PullTab, gml_VerTetBal, PullDeg, PullNul};" in s
        assert "#define   PullDegMax 32" in s


# --- the reference's example bodies, unmodified ------------------------------
def ref_body(path):
    return open(os.path.join(REF, path)).read()


needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")


@needs_ref
def test_reference_example_bodies_compile(mesh):
    crd, ref, tet = mesh
    tri = np.array([[0, 1, 2, 0], [1, 2, 3, 0]], np.int32)
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi, ti = g.set_mesh(crd, ref, tet)
        g.new_parameters(8, ref_body("MeshQuality/parameters.cl"))
        qal = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        opt = g.solution(G.GmlTetrahedra, 4, G.GmlFlt4, "OptCrd")
        res = g.solution(G.GmlVertices, 1, G.GmlFlt, "res")
        vdat = g.solution(G.GmlVertices, 1, G.GmlFlt4, "VerDat")
        tdat = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "TetDat")
        R, W, RW = G.GmlReadMode, G.GmlWriteMode, G.GmlReadMode | G.GmlWriteMode
        T, V = G.GmlTetrahedra, G.GmlVertices
        cases = [
            ("MeshQuality/compute_quality.cl", "compute_quality", T, [(vi, R, 0), (qal, W, 0)]),
            ("NodeSmoothing/quality.cl", "quality", T, [(vi, R, 0), (qal, W, 0)]),
            ("NodeSmoothing/scatter.cl", "scatter", T, [(vi, R, 0), (opt, W, 0)]),
            ("NodeSmoothing/gather.cl", "gather", V, [(res, W, 0), (opt, R | G.GmlVoyeurs, 0), (vi, RW, 0)]),
            ("GpuBandwidth/direct_vertices.cl", "direct_vertices", V, [(vi, R, 0), (vdat, W, 0)]),
            ("GpuBandwidth/direct_tets.cl", "direct_tets", T, [(ti, R, 0), (tdat, W, 0)]),
            ("GpuBandwidth/indirect_tets.cl", "indirect_tets", T, [(vi, R, 0), (tdat, W, 0)]),
            ("GpuBandwidth/vertices_ball.cl", "vertices_ball", V, [(vdat, W, 0), (tdat, R, 0)]),
        ]
        for path, name, typ, args in cases:
            k = g.compile(ref_body(path), name, typ, args)
            assert k > 0, path
        # TestFeatures: downlink with refs, uplink with voyeurs, FP64 data
        sol = g.solution(V, 2, G.GmlFlt4, "SolAtVer")
        mid = g.solution(T, 1, G.GmlFlt4, "TetMid")
        rv = g.solution(T, 1, G.GmlFlt, "ResVec")
        f64 = g.solution(T, 1, G.GmlDbl, "F64Dat")
        assert g.compile(ref_body("TestFeatures/downlink.cl"), "downlink", T,
                         [(vi, R | G.GmlRefFlag, 0), (sol, R, 0), (rv, W, 0), (mid, W, 0)])
        assert g.compile(ref_body("TestFeatures/uplink.cl"), "uplink", V, [(sol, W, 0), (mid, R | G.GmlVoyeurs, 0)])
        assert g.compile(ref_body("TestFeatures/double_precision.cl"), "double_precision", T, [(f64, R, 0)])


@needs_ref
def test_reference_ball_quality_bodies_compile():
    # BallQuality: triangles, a vertex->triangle uplink gathering the
    # triangle nodes and their quality
    crd = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [1, 1, 0, 0]], np.float32)
    tri = np.array([[0, 1, 2, 0], [1, 3, 2, 0]], np.int32)
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi = L.GmlNewMeshData(h, G.GmlVertices, 4)
        ti = L.GmlNewMeshData(h, G.GmlTriangles, 2)
        r = np.zeros(4, np.int32)
        assert L.GmlSetDataBlock(h, G.GmlVertices, 0, 3, G.ptr(crd), C.c_void_p(crd.ctypes.data + 48), G.ptr(r), G.ptr(r))
        assert L.GmlSetDataBlock(h, G.GmlTriangles, 0, 1, G.ptr(tri), C.c_void_p(tri.ctypes.data + 16), G.ptr(r), G.ptr(r))
        g.new_parameters(8, ref_body("BallQuality/parameters.cl"))
        tq = g.solution(G.GmlTriangles, 1, G.GmlFlt, "tq")
        vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
        assert g.compile(ref_body("BallQuality/triqal.cl"), "triqal", G.GmlTriangles, [(vi, G.GmlReadMode, 0), (tq, G.GmlWriteMode, 0)])
        assert g.compile(ref_body("BallQuality/verqal.cl"), "verqal", G.GmlVertices,
                         [(ti, G.GmlReadMode, 0), (tq, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)])
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTriangles, 4, 8)
        assert list(deg) == [1, 2, 2, 1]
        assert list(ball[1, :2]) == [(1 << 4) | 0, (0 << 4) | 1] and list(ball[2, :2]) == [(1 << 4) | 2, (0 << 4) | 2]


@needs_ref
def test_reference_heat_bodies_compile():
    # HeatCellCentered: triangle kernels over tri -> tet shells with TriRef,
    # tet kernels over the tet -> tri downlink, the parameter block of enums
    from gmlib_amd import heat as H
    crd, vref, tet, tri, side = H.heat_cube(2, 0.2, 1)
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi, ti = g.set_mesh(crd, vref, tet)
        g.new_parameters(76, ref_body("HeatCellCentered/param.cl"))
        nt = len(tri)
        assert L.GmlNewMeshData(h, G.GmlTriangles, nt)
        assert L.GmlSetDataBlock(h, G.GmlTriangles, 0, nt - 1, G.ptr(tri), C.c_void_p(tri.ctypes.data + (nt - 1) * 12),
                                 G.ptr(side), C.c_void_p(side.ctypes.data + (nt - 1) * 4))
        assert L.GmlExtractFaces(h) > nt
        idx = C.c_int(0)
        L.GmlGetMeshInfo(h, G.GmlTriangles, None, C.byref(idx))
        tr = idx.value
        T, F = G.GmlTetrahedra, G.GmlTriangles
        sol = g.solution(T, 1, G.GmlFlt, "SolTet")
        grd = g.solution(T, 1, G.GmlFlt4, "GrdTet")
        sx = g.solution(F, 1, G.GmlFlt, "SolExt")
        gx = g.solution(F, 1, G.GmlFlt4, "GrdExt")
        rhs = g.solution(T, 1, G.GmlFlt, "Rhs")
        dt = g.solution(T, 1, G.GmlFlt, "dt")
        R, W, RF = G.GmlReadMode, G.GmlWriteMode, G.GmlRefFlag
        cases = [("ini_tet", T, [(vi, R, 0), (sol, W, 0)]),
                 ("sol_ext", F, [(tr, R | RF, 0), (sol, R, 0), (sx, W, 0)]),
                 ("grd_tet", T, [(vi, R, 0), (sx, R, 0), (grd, W, 0)]),
                 ("grd_ext", F, [(tr, R | RF, 0), (grd, R, 0), (gx, W, 0)]),
                 ("flx_bal", T, [(vi, R, 0), (gx, R, 0), (rhs, W, 0)]),
                 ("tim_int", T, [(rhs, R, 0), (sol, R | W, 0)]),
                 ("dt", T, [(vi, R, 0), (dt, W, 0)])]
        for name, typ, args in cases:
            assert g.compile(ref_body("HeatCellCentered/%s.cl" % name), name, typ, args) > 0, name


def test_paired_form_emitted_only_where_it_is_safe(mesh):
    # element loops over downlinks get a paired form (two rows per thread);
    # kernels gathering through wide balls, bodies writing the parameter
    # block and bodies using work-item builtins keep the one-row form only
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        g.new_parameters(16, body("smooth_params"))
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        kq = g.compile(body("tet_meanratio"), "q", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        kg = g.compile(body("smooth_gather"), "gt", G.GmlVertices,
                       [(b, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
        kw = g.compile("qal = 1.f; GmlPar->omega += qal;", "w", G.GmlTetrahedra, [(q, G.GmlWriteMode, 0)])
        ki = g.compile("qal = (float)get_global_id(0);", "i", G.GmlTetrahedra, [(q, G.GmlWriteMode, 0)])
        kr = g.compile("if(TetIdx > 3) return; qal = 2.f;", "r", G.GmlTetrahedra, [(q, G.GmlWriteMode, 0)])
        src = g.L.GmlxGetKernelPairSource(g.h, kq).decode()
        assert g.L.GmlxGetKernelPairSource(g.h, kg) is None
        assert g.L.GmlxGetKernelPairSource(g.h, kw) is None
        assert g.L.GmlxGetKernelPairSource(g.h, ki) is None
        rsrc = g.L.GmlxGetKernelPairSource(g.h, kr).decode()
        assert g.L.GmlxGetKernelRows(g.h, kq) == 0 and g.L.GmlxGetKernelRows(g.h, kg) == 1
        assert g.L.GmlxSetKernelShape(g.h, kq, 128, 1, 2) == 1 and g.L.GmlxGetKernelRows(g.h, kq) == 2
        assert g.L.GmlxSetKernelShape(g.h, kg, 128, 1, 2) == 0
    # both rows read before either body; a row whose body returns early is not written
    assert src.index("gml_read(gml_r1);") < src.index("gml_body(gml_r0);")
    assert "if(gml_has1 && gml_r1.gml_done) gml_write(gml_r1);" in src
    assert "VerCrd[0] = VerCrdTab[ TetVer.s0 ];" in src
    assert "gml_x.gml_done = true;" in rsrc


def test_short_shells_gather_directly_and_pair(mesh):
    # a face's two cells (tri -> tet shell, degree <= 2) are gathered
    # directly, not staged in LDS, and the kernel gets a paired form; a body
    # that assigns to staged data falls back to the eager LDS reads
    from gmlib_amd import heat as H
    crd, vref, tet, tri, side = H.heat_cube(3, 0.2, 1)
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi, ti = g.set_mesh(crd, vref, tet)
        g.new_parameters(C.sizeof(H.HeatParams), H.body("heat_params"))
        nt = len(tri)
        assert L.GmlNewMeshData(h, G.GmlTriangles, nt)
        assert L.GmlSetDataBlock(h, G.GmlTriangles, 0, nt - 1, G.ptr(tri), C.c_void_p(tri.ctypes.data + (nt - 1) * 12),
                                 G.ptr(side), C.c_void_p(side.ctypes.data + (nt - 1) * 4))
        assert L.GmlExtractFaces(h) > nt
        idx = C.c_int(0)
        L.GmlGetMeshInfo(h, G.GmlTriangles, None, C.byref(idx))
        sol = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "SolTet")
        sx = g.solution(G.GmlTriangles, 1, G.GmlFlt, "SolExt")
        R = G.GmlReadMode
        k = g.compile(H.body("heat_face_value"), "fv", G.GmlTriangles,
                      [(idx.value, R | G.GmlRefFlag, 0), (sol, R, 0), (sx, G.GmlWriteMode, 0)])
        one, two = src_of(g, k), L.GmlxGetKernelPairSource(h, k)
        assert "gml_lds" not in one and two is not None
        two = two.decode()
        assert "SolTet[1] = (SolTetDeg >= 2) ? SolTetTab[ gml_TriTetBal[1] ] : SolTetNul;" in two
        assert two.index("gml_read(gml_r1);") < two.index("gml_body(gml_r0);")

        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        out = g.solution(G.GmlVertices, 1, G.GmlFlt4, "Out")
        ke = g.compile("Bary[0] = Bary[1] + Bary[2];\nOut = Bary[0];", "e", G.GmlVertices,
                       [(b, R, 0), (out, G.GmlWriteMode, 0)])
        se = src_of(g, ke)
    assert "gml_stg<" not in se and "gml_v_Bary[gml_u] = BaryTab[gml_x[gml_u]];" in se
    assert "Bary[0] = (BaryDeg >= 1) ? gml_s_Bary[ gml_l_VerTetBal[0] ] : BaryNul;" in se


def test_opencl_vector_literals_rewritten(mesh):
    # (T)(a, b, ...) of several parts is an OpenCL vector literal; a C++
    # compiler would read a comma expression.  The engine builds it with the
    # prelude's component-wise gml_vmake; a one-part (T)(x) stays a splat
    crd, ref, tet = mesh
    bod = """{
   float4 a = (float4)(1.f, 2.f, 3.f, 4.f);
   float2 h = (float2)(a.x, 0.5);
   float4 b = (float4)( h, (float2)(a.zw) );       /* vectors as parts, a nested literal */
   int4   n = (int4)(1, 2, (int2)(3, 4));
   float4 s = (float4)(0.25);
   const char *str = "(float4)(1, 2)";            /* strings are left alone */
   // (float4)(5, 6, 7, 8) in a comment too
   qal = dot(a + b + s, (float4)(n.x, n.y, n.z, n.w)) + (str[0] == '(' ? 0.f : 1.f);
}"""
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile(bod, "vlit", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        s = src_of(g, k)
    user = s[s.index("// USER'S KERNEL CODE"):]
    assert "gml_vmake<float4>(1.f, 2.f, 3.f, 4.f)" in user
    assert "gml_vmake<float2>(a.x, 0.5f)" in user
    assert "gml_vmake<float4>( h, (float2)(a.zw) )" in user
    assert "gml_vmake<int4>(1, 2, gml_vmake<int2>(3, 4))" in user
    assert "(float4)(0.25f)" in user
    assert '"(float4)(1, 2)"' in user and "// (float4)(5, 6, 7, 8) in a comment too" in user
    # a literal with the wrong number of components does not compile
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        tri = (C.c_int * 6)(vi, G.GmlReadMode, 0, q, G.GmlWriteMode, 0)
        assert g.L.GmlxCompileKernelv(g.h, b"{ float4 a = (float4)(1.f, 2.f, 3.f); qal = a.x; }", b"bad",
                                      G.GmlTetrahedra, 2, tri) == 0


def test_atomics_toolkit_and_work_item_bodies_compile(mesh):
    # OpenCL atomics on GmlManual data, a user toolkit of plain OpenCL-C
    # functions (compiled for host and device), and a body naming work-item
    # functions, whose launches keep the plain block order
    crd, ref, tet = mesh
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        c = g.solution(G.GmlVertices, 1, G.GmlInt, "cnt")
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile("{ for(int j = 0; j < 4; j++) { atomic_inc(&cntTab[TetVer[j]]); atom_max(&cntTab[0], j); } }",
                      "atomics", G.GmlTetrahedra, [(c, G.GmlReadMode | G.GmlWriteMode | G.GmlManual, 0)])
        assert not g.L.GmlxGetKernelPairSource(g.h, k) and not g.L.GmlxGetKernelTileSource(g.h, k)
        g.L.GmlIncludeUserToolkit(g.h, C.c_char_p(b"float twice(float x) { return 2.f * x; }\n"))
        k2 = g.compile("{ qal = twice(VerCrd[0].x) + (float)get_global_id(0); }", "tk_ids", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        assert "#pragma clang force_cuda_host_device begin" in src_of(g, k2)
        assert not g.L.GmlxSetKernelShape(g.h, k2, 128, 1, 1)
        assert g.L.GmlxSetKernelShape(g.h, k2, 128, 0, 1)
