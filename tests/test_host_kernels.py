"""The CPU baseline of the generated-kernel path (oracle/host_kernels.py):
the engine's generated smoothing kernels -- OpenCL-C prelude, GMlib
prologue/epilogue and the project's bodies -- compiled for x86-64 with the
host shim and run over the host cores.  Their sweeps must equal the oracle's
FP32 sweep (orc_smooth_sweep_f32) bit for bit on the same inputs: the same
operations in the same order, with FP contraction off on both sides.  This
also checks the generated prologue's ball indexing and padding on the CPU."""
import os

import numpy as np
import pytest

import gmlib_amd as G
from oracle import host_kernels as HK
from tests import oracle_lib as O

KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    src = HK.smoothing_sources(G, KDIR)
    so = HK.build(src, out=str(tmp_path_factory.mktemp("host")))
    return HK.HostKernels(so, src), src


def test_signature_and_loop(built):
    hk, src = built
    name, args = HK.kernel_signature(src["smooth_gather"])
    assert name == "smooth_gather"
    assert [n for t, n in args if "*" in t] == hk.args["smooth_gather"]
    assert [n for t, n in args][-2:] == ["count", "gml_remap"]
    assert HK.ball_width(src["smooth_gather"]) == 32
    assert hk.threads >= 1


def test_sell64_layout():
    deg = np.array([3, 0, 5] + [1] * 70, np.int32)
    ball = np.arange(len(deg) * 8, dtype=np.int32).reshape(len(deg), 8) + 100
    tab, off = HK.sell64(deg, ball)
    assert list(off) == [0, 5]                       # slice 0 as wide as degree 5, slice 1 as degree 1
    for v in range(len(deg)):
        for k in range(deg[v]):
            assert tab[off[v >> 6] * 64 + k * 64 + (v & 63)] == ball[v, k]


@pytest.mark.parametrize("n,sweeps", [(6, 3), (9, 5)])
def test_host_generated_sweeps_equal_oracle(built, n, sweeps):
    hk, _ = built
    crd, ref, tet = G.cube_mesh(n, 0.25, 7 + n)
    deg, ball, md = O.vertex_ball(tet, len(crd), 32)
    assert md <= 32
    c = np.array(crd, np.float32, copy=True)
    sm = HK.Smoothing(hk, c, ref, tet, deg, ball, 0.7)
    for _ in range(sweeps):
        sm.sweep()
    want = O.smooth32(np.ascontiguousarray(tet, np.int32), np.ascontiguousarray(ref, np.int32), deg, ball,
                      crd, sweeps, 0.7)
    assert np.array_equal(c[:, :3], want[:, :3])
    assert not np.array_equal(c[:, :3], crd[:, :3])        # interior vertices moved
    assert np.array_equal(c[ref != 0, :3], crd[ref != 0, :3])  # boundary fixed


def test_staged_source_is_refused():
    crd, ref, tet = G.cube_mesh(4, 0.2, 3, True)
    with G.Gml(host_only=True) as g:
        G.lib().GmlxSetStaging(g.h, 1)
        vi, ti = g.set_mesh(crd, ref, tet)
        g.new_parameters(16, open(os.path.join(KDIR, "smooth_params.cl")).read())
        b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        k = g.compile(open(os.path.join(KDIR, "smooth_gather.cl")).read(), "smooth_gather", G.GmlVertices,
                      [(b, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
        s = g.L.GmlxGetKernelSource(g.h, k).decode()
    with pytest.raises(ValueError):
        HK.wrapper(s)


def test_host_generated_quality_and_ball_min_equal_oracle(tmp_path):
    src = HK.quality_sources(G, KDIR)
    hk = HK.HostKernels(HK.build(src, out=str(tmp_path), tag="quality"), src)
    assert HK.ball_width(src["ball_min"], "qal") == 32
    crd, ref, tet = G.cube_mesh(12, 0.3, 5)
    deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
    hq = HK.Quality(hk, crd, tet, deg, ball)
    hq.step()
    want = O.tet_quality32(tet, crd)
    assert np.array_equal(hq.qal, want)
    assert np.array_equal(hq.vq, O.ball_min(deg, ball, want))
