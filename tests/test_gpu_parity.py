"""GPU parity of the generated-kernel path on an MI355X, against the oracle.

Tolerances (from the workload definition): topology, ball lists and min/max
bit-exact; FP32 quality values and sums within 1e-5 relative of the FP64
oracle; smoothed coordinates within 1e-5 absolute (unit cube) of the FP64
oracle.  Each test prints its measured error against the FP64 oracle.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


def quality_setup(g, crd, ref, tet):
    vi, ti = g.set_mesh(crd, ref, tet)
    q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
    k = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra,
                  [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
    return vi, ti, q, k


@pytest.mark.parametrize("n,jit,seed", [(20, 0.2, 1), (7, 0.3, 2), (1, 0.0, 3)])
def test_tet_quality_and_reductions(n, jit, seed):
    crd, ref, tet = G.cube_mesh(n, jit, seed)
    print("case: quality n=%d jitter=%g seed=%d" % (n, jit, seed), flush=True)
    with G.Gml() as g:
        vi, ti, q, k = quality_setup(g, crd, ref, tet)
        for _ in range(9):          # past the launch-shape tuner
            g.launch(k)
        qg = g.download(q, np.float32, (len(tet),))
        q64 = O.tet_quality64(tet, crd)
        err = np.abs(qg - q64) / np.maximum(np.abs(q64), 1e-3)
        print("quality max rel err vs FP64 oracle: %.3e" % err.max())
        assert err.max() <= 1e-5
        # reductions: min/max bit-exact vs the same FP32 array, sums vs FP64
        assert g.reduce(q, G.GmlMin) == float(qg.min())
        assert g.reduce(q, G.GmlMax) == float(qg.max())
        assert g.reduce(q, G.GmlLinf) == float(np.abs(qg).max())
        assert g.reduce(q, G.GmlL0) == float(np.count_nonzero(qg))
        s = g.reduce(q, G.GmlSum)
        exact = math.fsum(qg.astype(np.float64))
        assert abs(s - exact) <= 1e-12 * len(qg)
        print("sum: gpu %.17g fsum(gpu array) %.17g fp64-oracle %.17g" % (s, exact, math.fsum(q64)))
        assert abs(s - math.fsum(q64)) <= 1e-5 * abs(math.fsum(q64))
        assert g.reduce(q, G.GmlL1) == pytest.approx(math.fsum(np.abs(qg.astype(np.float64))), rel=1e-13)
        assert g.reduce(q, G.GmlL2) == pytest.approx(math.fsum(qg.astype(np.float64) ** 2), rel=1e-13)
        assert g.kernel_time(k) > 0


def test_gpu_vertex_ball_bit_exact():
    crd, ref, tet = G.cube_mesh(20, 0.2, 1)
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        assert g.L.GmlxBuildLink(g.h, G.GmlVertices, G.GmlTetrahedra)
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, len(crd), 32)
        assert g.link_info(G.GmlVertices, G.GmlTetrahedra) == (len(crd), 32, 0, 0)
    d0, b0, md = O.vertex_ball(tet, len(crd), 32)
    assert md == 24
    assert np.array_equal(deg, d0)
    assert np.array_equal(ball, b0)


def test_gpu_ball_random_soup_and_high_degree():
    # random tets over few vertices: degrees above 32 give high rows
    rng = np.random.default_rng(5)
    nv = 64
    tet = np.array([rng.choice(nv, 4, replace=False) for _ in range(700)], np.int32)
    crd = np.zeros((nv, 4), np.float32)
    ref = np.zeros(nv, np.int32)
    with G.Gml() as g:
        g.set_mesh(crd, ref, tet)
        assert g.L.GmlxBuildLink(g.h, G.GmlVertices, G.GmlTetrahedra)
        d0, b0, md = O.vertex_ball(tet, nv, 128)
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, nv, 128)
        info = g.link_info(G.GmlVertices, G.GmlTetrahedra)
    assert md > 32 and info[1] == 32 and info[3] >= md
    first = int(np.flatnonzero(d0 > 32)[0])
    assert info[0] == first and info[2] == nv - first
    assert np.array_equal(deg, d0) and np.array_equal(ball, b0)


def test_ball_min_generated():
    crd, ref, tet = G.cube_mesh(20, 0.2, 1)
    with G.Gml() as g:
        vi, ti, q, kq = quality_setup(g, crd, ref, tet)
        vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vq")
        kb = g.compile(body("ball_min"), "ball_min", G.GmlVertices, [(q, G.GmlReadMode, 0), (vq, G.GmlWriteMode, 0)])
        g.launch(kq)
        for _ in range(8):
            g.launch(kb)
        qg = g.download(q, np.float32, (len(tet),))
        vg = g.download(vq, np.float32, (len(crd),))
        d0, b0, _ = O.vertex_ball(tet, len(crd), 32)
        assert np.array_equal(vg, O.ball_min(d0, b0, qg))
        assert g.reduce(vq, G.GmlMin) == float(qg.min())


def smoothing_setup(g, crd, ref, tet, omega):
    vi, ti = g.set_mesh(crd, ref, tet)
    par = g.new_parameters(16, body("smooth_params"))
    C.cast(par, C.POINTER(C.c_float))[0] = omega
    assert g.L.GmlUploadParameters(g.h) == 1
    b = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
    ks = g.compile(body("smooth_scatter"), "smooth_scatter", G.GmlTetrahedra,
                   [(vi, G.GmlReadMode, 0), (b, G.GmlWriteMode, 0)])
    kg = g.compile(body("smooth_gather"), "smooth_gather", G.GmlVertices,
                   [(b, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
    return vi, b, ks, kg


@pytest.mark.parametrize("n,sweeps", [(20, 10), (9, 3)])
def test_smoothing_sweeps_generated(n, sweeps):
    crd, ref, tet = G.cube_mesh(n, 0.2, 7)
    omega = 0.2
    with G.Gml() as g:
        vi, b, ks, kg = smoothing_setup(g, crd, ref, tet, omega)
        for _ in range(sweeps):
            g.launch(ks)
            g.launch(kg)
        cg, rg = g.vertices(len(crd))
    d0, b0, _ = O.vertex_ball(tet, len(crd), 32)
    c32 = O.smooth32(tet, ref, d0, b0, crd, sweeps, omega)
    c64 = O.smooth64(tet, ref, d0, b0, crd, sweeps, omega)
    e32 = np.abs(cg[:, :3] - c32[:, :3]).max()
    e64 = np.abs(cg[:, :3] - c64).max()
    print("smoothing n=%d sweeps=%d: max |gpu - oracle f32| %.3e, |gpu - oracle f64| %.3e" % (n, sweeps, e32, e64))
    assert e64 <= 1e-5 and e32 <= 1e-5
    assert np.array_equal(cg[ref != 0], crd[ref != 0])
    assert np.array_equal(rg, ref)


# Two bodies wired the way the reference's NodeSmoothing example wires its
# scatter and gather kernels (NodeSmoothing.c:137-141): a tet body writing
# four float4 items per tet, a vertex body reading them through the vertex
# -> tet ball with voyeurs (each ball entry's local index selects its item),
# looping to ...DegMax over the ...Nul-padded ball, testing ...Deg, and
# relaxing with mix().  The scheme (volume-weighted pull towards the
# centroids of the opposite faces) is the oracle's orc_face_pull_sweep_f64.
FACE_PULL_SCATTER = """{
   float4 p0 = VerCrd[0], p1 = VerCrd[1], p2 = VerCrd[2], p3 = VerCrd[3];
   float4 sum = p0 + p1 + p2 + p3;
   float  w = fabs(dot(cross(p1 - p0, p2 - p0), p3 - p0)) * (1.f / 6.f);
   for(int j = 0; j < 4; j++)
   {
      float4 f = (sum - VerCrd[j]) * (1.f / 3.f);
      f.w = w;
      Pull[j] = f;
   }
}"""

FACE_PULL_GATHER = """{
   float4 acc = (float4)(0.f);
   float  wsum = 0.f;
   for(int k = 0; k < PullDegMax; k++)
   {
      float4 f = Pull[k][VerTetVoy[k]];      /* PullNul (zeros) past the degree */
      acc += f * (float4)(f.w);
      wsum += f.w;
   }
   move = 0.f;
   if(PullDeg > 0)
   {
      float4 target = acc / (float4)(wsum);
      target.w = 0.f;
      float4 next = mix(VerCrd, target, 0.25f);
      move = dot(next - VerCrd, next - VerCrd);
      VerCrd = next;
   }
}"""


@pytest.mark.parametrize("n,sweeps,isolated", [(12, 3, False), (7, 2, True)])
def test_face_pull_uplink_pattern_generated(n, sweeps, isolated):
    crd, ref, tet = G.cube_mesh(n, 0.2, 4)
    if isolated:
        # a vertex that no tet uses: its ball is empty (PullDeg == 0)
        crd = np.vstack([crd, [[0.5, 0.5, 2.0, 0.0]]]).astype(np.float32)
        ref = np.append(ref, 0).astype(np.int32)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        pull = g.solution(G.GmlTetrahedra, 4, G.GmlFlt4, "Pull")
        move = g.solution(G.GmlVertices, 1, G.GmlFlt, "move")
        ks = g.compile(FACE_PULL_SCATTER, "face_pull_scatter", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (pull, G.GmlWriteMode, 0)])
        kg = g.compile(FACE_PULL_GATHER, "face_pull_gather", G.GmlVertices,
                       [(move, G.GmlWriteMode, 0), (pull, G.GmlReadMode | G.GmlVoyeurs, 0),
                        (vi, G.GmlReadMode | G.GmlWriteMode, 0)])
        src = g.L.GmlxGetKernelSource(g.h, kg).decode()
        assert "gml_upn<float4, 4, 32>" in src and "VerTetVoy" in src
        for _ in range(sweeps):
            g.launch(ks)
            g.launch(kg)
        cg, _ = g.vertices(len(crd))
        mv = g.download_raw(move, np.float32, (len(crd),))
        msum = g.reduce(move, G.GmlSum)
    d0, b0, _ = O.vertex_ball(tet, len(crd), 32)
    c64, _, r64 = O.face_pull64(tet, d0, b0, crd, sweeps)
    e = np.abs(cg[:, :3] - c64).max()
    er = np.abs(mv - r64).max() / r64.max()
    print("face pull n=%d: max |gpu - fp64 oracle| = %.3e, residual max rel %.3e, sum gpu %.6e fp64 %.6e"
          % (n, e, er, msum, r64.sum()))
    assert e <= 1e-5
    assert er <= 1e-4
    assert msum == pytest.approx(r64.sum(), rel=1e-4)
    if isolated:
        assert np.array_equal(cg[-1], crd[-1]) and mv[-1] == 0.0


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1025, 1 << 20, (1 << 22) + 3])
def test_reductions_sizes(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n).astype(np.float32)
    x[::7] = 0.
    with G.Gml() as g:
        g.L.GmlNewMeshData(g.h, G.GmlVertices, n)
        d = g.solution(G.GmlVertices, 1, G.GmlFlt, "x")
        nb = C.c_size_t(0)
        dp = g.L.GmlxGetDevicePointer(g.h, d, C.byref(nb))
        g.L.GmlxSync(g.h)          # the zero fill runs on the library's stream
        G._hip_h2d(dp, x)
        assert g.reduce(d, G.GmlMin) == float(x.min())
        assert g.reduce(d, G.GmlMax) == float(x.max())
        assert g.reduce(d, G.GmlLinf) == float(np.abs(x).max())
        assert g.reduce(d, G.GmlL0) == float(np.count_nonzero(x))
        x64 = x.astype(np.float64)
        for op, ref in ((G.GmlSum, math.fsum(x64)), (G.GmlL1, math.fsum(np.abs(x64))), (G.GmlL2, math.fsum(x64 * x64))):
            got = g.reduce(d, op)
            assert abs(got - ref) <= 1e-13 * max(1.0, math.fsum(np.abs(x64)) if op != G.GmlL2 else ref), op
        # the same vector twice in a row, then another vector (reference Q3)
        assert g.reduce(d, G.GmlSum) == g.reduce(d, G.GmlSum)


def test_set_get_lines_on_device():
    with G.Gml() as g:
        vi = g.L.GmlNewMeshData(g.h, G.GmlVertices, 5)
        for i in range(5):
            assert G.set_line(g.h, vi, i, C.c_double(i), C.c_double(i + .25), C.c_double(-i), C.c_int(i % 2))
        crd, ref = g.vertices(5)
        assert np.array_equal(crd[:, 1], np.arange(5) + .25) and np.array_equal(ref, [0, 1, 0, 1, 0])
        x = C.c_double()
        assert G.get_line(g.h, vi, 3, C.byref(x), C.byref(C.c_double()), C.byref(C.c_double())) and x.value == 3.0


def test_c_user_program_runs():
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = os.path.join(tempfile.mkdtemp(), "mesh_quality")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(root, "include"),
                    os.path.join(root, "gmlib_amd", "examples", "mesh_quality.c"),
                    "-L", os.path.join(root, "gmlib_amd"), "-lgml", "-Wl,-rpath," + os.path.join(root, "gmlib_amd"),
                    "-o", out], check=True)
    r = subprocess.run([out, "6"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    crd, ref, tet = G.cube_mesh(6, 0.2, 1)
    q64 = O.tet_quality64(tet, crd)
    got = float(r.stdout.split("min quality ")[1].split(",")[0])
    assert abs(got - q64.min()) <= 1e-5 * abs(q64.min())
