"""Device-handle paths of the data lifecycle that the host-only tests cannot
reach: GmlExportSolution of a field a kernel has just written (the device
copy must be brought back first, as GmlGetDataLine's lazy download does,
sources/gmlib3.c:1921-1922 and 4606-4740), GmlFreeData of device data and
the reuse of its slot (gmlib3.c:1779-1805), GmlGetFlops / GmlGetMemoryUsage
after launches, and GmlListGPU."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O
from tests.test_meshb import GMF, read_sol_binary

pytestmark = pytest.mark.gpu

MEANRATIO = """
{
   float4 e0 = VerCrd[1] - VerCrd[0], e1 = VerCrd[2] - VerCrd[0], e2 = VerCrd[3] - VerCrd[0];
   float4 e3 = VerCrd[2] - VerCrd[1], e4 = VerCrd[3] - VerCrd[1], e5 = VerCrd[3] - VerCrd[2];
   float  s  = dot(e0, e0) + dot(e1, e1) + dot(e2, e2) + dot(e3, e3) + dot(e4, e4) + dot(e5, e5);
   float  v  = dot(cross(e0, e1), e2) * (1.f / 6.f);
   qal = 12.f * cbrt(9.f * v * v) / s;
   if(v < 0.f)
      qal = -qal;
}
"""


def test_export_of_a_kernel_written_field(tmp_path):
    crd, ref, tet = G.cube_mesh(5, 0.2, 9, True)
    T = len(tet)
    with G.Gml(0) as g:
        L, h = g.L, g.h
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile(MEANRATIO, "quality_for_export", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        g.launch(k)
        p = str(tmp_path / "q.solb")
        assert L.GmlExportSolution(C.c_size_t(h), C.c_char_p(p.encode()), C.c_int(q), C.c_int(0)) == 1
        dev = g.download(q, np.float32, (T,))
    typ, vals = read_sol_binary(open(p, "rb").read())[GMF["SolAtTetrahedra"]]
    assert typ == (1,) and vals.shape == (T, 1)
    # the file holds what the kernel wrote, bit for bit, and that is the
    # mean ratio of the mesh within FP32 rounding of the FP64 oracle
    assert np.array_equal(vals[:, 0], dev)
    q64 = O.tet_quality64(tet, crd)
    assert np.max(np.abs(dev - q64) / np.maximum(np.abs(q64), 1e-3)) <= 1e-5


def test_free_device_data_and_reuse_the_slot():
    crd, ref, tet = G.cube_mesh(4, 0.2, 3, True)
    T = len(tet)
    with G.Gml(0) as g:
        L, h = g.L, g.h
        vi, ti = g.set_mesh(crd, ref, tet)
        used0 = L.GmlGetMemoryUsage(h)
        a = g.solution(G.GmlTetrahedra, 4, G.GmlFlt4, "big")
        used1 = L.GmlGetMemoryUsage(h)
        assert used1 - used0 >= T * 64
        assert L.GmlFreeData(h, a) == 1
        assert L.GmlFreeData(h, a) == 0              # already free
        assert L.GmlGetMemoryUsage(h) == used0
        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        assert b == a                                 # the lowest free slot is reused
        k = g.compile(MEANRATIO, "quality_after_free", G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (b, G.GmlWriteMode, 0)])
        assert L.GmlLaunchKernel(h, k) == 1
        qg = g.download(b, np.float32, (T,))
        assert L.GmlGetFlops(h) >= 0.
    q64 = O.tet_quality64(tet, crd)
    assert np.max(np.abs(qg - q64) / np.maximum(np.abs(q64), 1e-3)) <= 1e-5


def test_list_gpu_runs():
    G.lib().GmlListGPU()
