"""FP64 user data through generated kernels on the GPU (GmlDbl solution
fields, GmlCheckFP64; the reference's TestFeatures example runs a body on a
GmlDbl tet field, examples/TestFeatures/TestFeatures.c:126-140): a vertex
GmlDbl field gathered through the tet -> vertex downlink in every launch
form, a tet GmlDbl field gathered through the vertex -> tet ball, a GmlDbl4
field built from a vector literal, and line transfers of double data.  The
values differ from one another only below FP32 resolution, so a result that
went through FP32 anywhere fails; the expected values are numpy FP64 sums in
the same order, compared exactly."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_fp64_fields_through_downlink_uplink_and_literals():
    crd, ref, tet = G.cube_mesh(14, 0.2, 12)
    V, T = len(crd), len(tet)
    vd0 = 1.0 + np.arange(V, dtype=np.float64) * 2.0 ** -40          # all equal in FP32
    with G.Gml() as g:
        L, h = g.L, g.h
        assert L.GmlCheckFP64(h) == 1
        vi, ti = g.set_mesh(crd, ref, tet)
        vd = g.solution(G.GmlVertices, 1, G.GmlDbl, "vd")
        td = g.solution(G.GmlTetrahedra, 1, G.GmlDbl, "td")
        vs = g.solution(G.GmlVertices, 1, G.GmlDbl, "vs")
        t4 = g.solution(G.GmlTetrahedra, 1, G.GmlDbl4, "t4")
        g.upload(vd, vd0)
        # tet: mean of its vertices' doubles (the element-tile form stages them as doubles)
        kt = g.compile("{ td = (vd[0] + vd[1] + vd[2] + vd[3]) * 0.25; }", "tet_mean", G.GmlTetrahedra,
                       [(vd, G.GmlReadMode, 0), (td, G.GmlWriteMode, 0)])
        # vertex: sum of its ball's tet doubles, to the padded width (tdNul = 0)
        kv = g.compile("{ vs = 0.; for(int i = 0; i < tdDegMax; i++) vs += td[i]; }", "ball_sum", G.GmlVertices,
                       [(td, G.GmlReadMode, 0), (vs, G.GmlWriteMode, 0)])
        # a double4 vector literal of several parts
        k4 = g.compile("{ t4 = (double4)((double2)(td, td * 2.), (double)TetIdx, 0.); }", "tet_d4", G.GmlTetrahedra,
                       [(td, G.GmlReadMode, 0), (t4, G.GmlWriteMode, 0)])
        assert L.GmlxGetKernelTileSource(h, kt), "the double gather has an element-tile form"
        outs = []
        for rows in (1, 2, -1):
            assert L.GmlxSetKernelShape(h, kt, 128, 1, rows)
            g.launch(kt)
            outs.append(g.download_raw(td, np.float64, (T,)))
        g.launch(kv)
        g.launch(k4)
        tdg = outs[0]
        vsg = g.download_raw(vs, np.float64, (V,))
        t4g = g.download_raw(t4, np.float64, (T, 4))
        # line transfers of double data (the reference's lazy rule: reading
        # line 0 downloads, setting the last line uploads the mirror)
        one = C.c_double()
        assert G.get_line(h, td, 0, C.byref(one)) and one.value == tdg[0]
        assert G.get_line(h, td, 7, C.byref(one)) and one.value == tdg[7]
        x = C.c_double(3.0 + 2.0 ** -45)
        assert G.set_line(h, td, T - 1, C.byref(x))
        assert G.get_line(h, td, 0, C.byref(one)) and one.value == tdg[0]
        assert G.get_line(h, td, T - 1, C.byref(one)) and one.value == x.value
        assert g.download_raw(td, np.float64, (T,))[T - 1] == x.value
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, V, 32)
    for o in outs[1:]:
        assert np.array_equal(o, tdg)
    want = (((vd0[tet[:, 0]] + vd0[tet[:, 1]]) + vd0[tet[:, 2]]) + vd0[tet[:, 3]]) * 0.25
    assert np.array_equal(tdg, want)
    assert len(np.unique(tdg.astype(np.float32))) < len(np.unique(tdg)) // 100   # FP32 could not tell them apart
    # the ball sums in ball order (descending tets), padding slots adding 0
    s = np.zeros(V)
    for k in range(32):
        e = ball[:, k] >> 4
        s = s + np.where(k < deg, want[e], 0.0)
    assert np.array_equal(vsg, s)
    assert np.array_equal(t4g[:, 0], tdg) and np.array_equal(t4g[:, 1], tdg * 2.0)
    assert np.array_equal(t4g[:, 2], np.arange(T, dtype=np.float64)) and not t4g[:, 3].any()
