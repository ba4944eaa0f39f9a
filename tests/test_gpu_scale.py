"""GPU parity at the full workload sizes, against the oracle:

- 256^3 Kuhn cube (16,974,593 vertices, 100,663,296 tets): 10 Jacobi
  smoothing sweeps through the drop-in generated kernels, the exact fused
  sweep and the operator-form sweep, then tet quality and its minimum;
- the 3x3 edge Laplacian of the 128^3 cube (2,146,689 block rows, 31.8M
  blocks): SpMV, FP32 CG and FP64-refined CG, residuals recomputed by the
  oracle's FP64 product; the SpMV of the 256^3 cube's edge Laplacian
  (16,974,593 block rows, 253M blocks);
- 512 cube meshes with cells per side log-uniform in [8, 96] (342M tets)
  in one batch: per-tet quality, per-vertex ball minimum and per-mesh
  minimum / mean;
- a served mix on meshes of 10k to 20M tets, every result against the
  oracle on the same mesh.
"""
import ctypes as C
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
KDIR = __import__("os").path.join(__import__("os").path.dirname(G.__file__), "kernels")
EPS32 = 2.0 ** -24


def body(name):
    return open(KDIR + "/" + name + ".cl").read()


def log(*a):
    print(*a, flush=True)


def quality_ok(q, crd3, tet):
    """FP32 quality against FP64 on the same coordinates: within
    8 eps (1 + |q|^-1/2) everywhere (the volume's FP32 cancellation grows as
    |q|^-3/2 on nearly flat tets) and 1e-5 relative where |q| >= 0.05."""
    crd4 = np.zeros((len(crd3), 4), np.float32)
    crd4[:, :3] = crd3
    q64 = O.tet_quality64(tet, crd4)
    ae = np.abs(q - q64)
    scaled = float((ae / (EPS32 * (1 + 1 / np.sqrt(np.maximum(np.abs(q64), 1e-12))))).max())
    big = np.abs(q64) >= 0.05
    rel = float((ae[big] / np.abs(q64[big])).max())
    return scaled, rel


def test_smoothing_256cube_three_forms_vs_oracle():
    n, sweeps, omega = 256, 10, 0.2
    t0 = time.time()
    crd, ref, tet = G.cube_mesh(n, 0.2, 2024, True)
    V, T = len(crd), len(tet)
    assert (V, T) == (16974593, 100663296)
    crd0 = np.ascontiguousarray(crd)
    log("mesh %d^3 in %.1fs" % (n, time.time() - t0))
    out = {}
    with G.Gml() as g:
        L, h = g.L, g.h
        vi, ti = g.set_mesh(crd, ref, tet)
        par = g.new_parameters(16, body("smooth_params"))
        C.cast(par, C.POINTER(C.c_float))[0] = omega
        L.GmlUploadParameters(h)
        bary = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "Bary")
        qal = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        ks = g.compile(body("smooth_scatter"), "smooth_scatter", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (bary, G.GmlWriteMode, 0)])
        kg = g.compile(body("smooth_gather"), "smooth_gather", G.GmlVertices,
                       [(bary, G.GmlReadMode, 0), (vi, G.GmlReadMode | G.GmlWriteMode | G.GmlRefFlag, 0)])
        kq = g.compile(body("tet_meanratio"), "tet_meanratio", G.GmlTetrahedra,
                       [(vi, G.GmlReadMode, 0), (qal, G.GmlWriteMode, 0)])

        def restore():
            assert L.GmlSetDataBlock(h, G.GmlVertices, 0, V - 1, G.ptr(crd0),
                                     C.c_void_p(crd0.ctypes.data + (V - 1) * 16), G.ptr(ref),
                                     C.c_void_p(ref.ctypes.data + (V - 1) * 4))

        for form in ("generated", "fused", "operator"):
            restore()
            if form == "generated":
                for _ in range(sweeps):
                    g.launch(ks)
                    g.launch(kg)
                g.launch(kq)
            else:
                assert L.GmlxSetSmoothMode(h, 1 if form == "fused" else 2)
                assert L.GmlxSmoothSweeps(h, bary, sweeps, omega) == 1
                assert L.GmlxTetQuality(h, qal) == 1
            qmin = g.reduce(qal, G.GmlMin)
            c, _ = g.vertices(V)
            q = g.download_raw(qal, np.float32, (T,))
            out[form] = (c[:, :3].copy(), q, qmin)
            log("%s: min quality %.7f" % (form, qmin))
        deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, V, 32)
    # the exact fused sweep is the drop-in sweep in another kernel shape
    assert np.array_equal(out["generated"][0], out["fused"][0])
    c32 = O.smooth32(tet, ref, deg, ball, crd0, sweeps, omega)[:, :3]
    c64 = O.smooth64(tet, ref, deg, ball, crd0, sweeps, omega)
    for form, (c, q, qmin) in out.items():
        d32 = float(np.abs(c - c32).max())
        d64 = float(np.abs(c - c64).max())
        scaled, rel = quality_ok(q, c, tet)
        log("%s: max |x - oracle f32| %.2e, |x - oracle f64| %.2e; quality %.2f eps-scaled, %.2e relative"
            % (form, d32, d64, scaled, rel))
        assert d32 <= 1e-5 and d64 <= 1e-5
        assert qmin == float(q.min())                  # bit-exact on the same FP32 array
        assert scaled <= 8 and rel <= 1e-5
    # boundary vertices stayed where they were
    fixed = ref != 0
    for form, (c, _, _) in out.items():
        assert np.array_equal(c[fixed], crd0[fixed, :3])


def test_spmv_cg_128cube_vs_oracle():
    n = 128
    crd, ref, tet = G.cube_mesh(n, 0.2, 2024, True)
    V, T = len(crd), len(tet)
    L = G.lib()
    rowptr = np.empty(V + 1, np.int32)
    nb = L.GmlxEdgeLaplacian3(V, T, G.ptr(tet), 2024, 0.01, G.ptr(rowptr), None, None, 0)
    assert nb == V + 2 * 14827904
    col = np.empty(nb, np.int32)
    val = np.empty(nb * 9, np.float32)
    assert L.GmlxEdgeLaplacian3(V, T, G.ptr(tet), 2024, 0.01, G.ptr(rowptr), G.ptr(col), G.ptr(val), nb) == nb
    rhs = np.random.default_rng(7).normal(size=V * 3).astype(np.float32)
    b64 = rhs.astype(np.float64)
    bn = float(np.linalg.norm(b64))
    with G.Gml() as g:
        m = L.GmlNewMatrix(g.h, V, nb, 3, G.ptr(val), G.ptr(col), G.ptr(rowptr), G.GmlFlt)
        vx = L.GmlNewVector(g.h, V, 3, G.ptr(rhs), G.GmlFlt)
        vy = L.GmlNewVector(g.h, V, 3, None, G.GmlFlt)
        vs = L.GmlNewVector(g.h, V, 3, None, G.GmlDbl)
        assert min(m, vx, vy, vs) > 0
        assert L.GmlMultMatVec(g.h, m, vy, vx) == 1
        y = np.empty(V * 3, np.float32)
        assert L.GmlxGetVector(g.h, vy, G.ptr(y))
        y64 = O.bsr_spmv64(rowptr, col, val, b64, 3)
        err = float(np.abs(y - y64).max() / np.abs(y64).max())
        log("SpMV: max |y - y64| / max |y64| = %.2e" % err)
        assert err <= 1e-5
        for tol, refine, bound in ((1e-5, 0, 2e-5), (1e-10, 1, 1e-10)):
            rr, it = C.c_double(), C.c_int()
            assert L.GmlxSolveCG(g.h, m, vx, vs, tol, 20000, refine, C.byref(rr), C.byref(it)) == 1
            x = np.empty(V * 3, np.float64)
            assert L.GmlxGetVector(g.h, vs, G.ptr(x))
            res = float(np.linalg.norm(b64 - O.bsr_spmv64(rowptr, col, val, x, 3)) / bn)
            log("CG tol %.0e refine %d: %d iterations, GPU FP64 residual %.3e, oracle %.3e"
                % (tol, refine, it.value, rr.value, res))
            assert res <= bound and it.value > 0
            assert abs(res - rr.value) <= 1e-3 * res + 1e-14


def test_batch_512_meshes_vs_oracle():
    rng = np.random.default_rng(2024)
    sizes = np.round(np.exp(rng.uniform(np.log(8), np.log(96), 512))).astype(int)
    with ThreadPoolExecutor(16) as ex:
        ms = list(ex.map(lambda i: G.cube_mesh(int(sizes[i]), 0.3, 7000 + i), range(len(sizes))))
    nver, ntet, crd, tet = G.concat_meshes(ms)
    V, T, M = int(nver.sum()), int(ntet.sum()), len(ms)
    log("batch: %d meshes, V=%d T=%d" % (M, V, T))
    with G.Gml() as g:
        b = g.L.GmlxNewBatch(g.h, M, G.ptr(nver), G.ptr(ntet), G.ptr(crd), G.ptr(tet))
        assert b > 0
        tv, pad = C.c_int(), C.c_int()
        assert g.L.GmlxBatchTune(g.h, b, C.byref(tv), C.byref(pad)) == 1
        mn = np.empty(M, np.float32)
        mean = np.empty(M, np.float64)
        assert g.L.GmlxBatchQuality(g.h, b, G.ptr(mn), G.ptr(mean)) == 1
        tq = np.empty(T, np.float32)
        vm = np.empty(V, np.float32)
        assert g.L.GmlxGetBatchField(g.h, b, G.ptr(tq), G.ptr(vm)) == 1
    del crd, tet
    vo = to = 0
    worst_scaled = worst_rel = 0.
    for i, (crd_m, ref_m, tet_m) in enumerate(ms):
        q = tq[to:to + len(tet_m)]
        v = vm[vo:vo + len(crd_m)]
        scaled, rel = quality_ok(q, crd_m[:, :3], tet_m)
        worst_scaled, worst_rel = max(worst_scaled, scaled), max(worst_rel, rel)
        deg, ball, md = O.vertex_ball(tet_m, len(crd_m), 32)
        assert md <= 32
        assert np.array_equal(v, O.ball_min(deg, ball, q)), "mesh %d: ball minimum" % i
        assert mn[i] == q.min(), "mesh %d: minimum" % i
        assert mean[i] == pytest.approx(math.fsum(q.astype(np.float64)) / len(q), rel=1e-12)
        vo += len(crd_m)
        to += len(tet_m)
    log("TV %d, Pad %d; quality %.2f eps-scaled, %.2e relative (|q| >= 0.05)"
        % (tv.value, pad.value, worst_scaled, worst_rel))
    assert worst_scaled <= 8 and worst_rel <= 1e-5


def test_served_mix_10k_to_20m_tets_vs_oracle():
    """Quality, 20-sweep smoothing and CG jobs on registered meshes of 10k to
    20M tets, submitted at once to a 4-slot server: quality min and mean
    within 1e-5 of the FP64 oracle; smoothing against the oracle's FP32
    sweeps (the served sweep is the operator form: equal up to FP32 rounding
    of the coordinates, which moves the minimum by about that rounding over
    the edge length, bound 1e-4; mean 1e-5); CG at its FP64 tolerance."""
    L = G.lib()
    cells = [12, 30, 60, 100, 149]                      # 10,368 to 19,844,094 tets
    ms = [G.cube_mesh(n, 0.2, 500 + n) for n in cells]
    srv = L.GmlxServeStart(0, 64.0, 4, 1)
    keys = [L.GmlxServeAddMesh(srv, len(c), len(t), G.ptr(c), G.ptr(r), G.ptr(t)) for c, r, t in ms]
    assert all(k > 0 for k in keys)
    jobs = [(t, i, p) for i in range(len(ms))
            for t, p in ((G.GmlxJobQuality, 0.0), (G.GmlxJobSmooth, 20.0), (G.GmlxJobCG, 1e-8))]
    ids = [L.GmlxServeSubmitMesh(srv, t, keys[i], p) for t, i, p in jobs]
    res = []
    for j in ids:
        r, tim = (C.c_double * 4)(), (C.c_double * 4)()
        assert L.GmlxServeWait(srv, j, r, tim) == 1
        res.append(tuple(r))
    info = (C.c_double * 8)()
    assert L.GmlxServeInfo(srv, info)
    L.GmlxServeStop(srv)
    log("served %d jobs, up to %d at once" % (len(jobs), info[5]))
    for (t, i, p), r in zip(jobs, res):
        crd, ref, tet = ms[i]
        if t == G.GmlxJobCG:
            assert r[0] <= 1e-8 and r[1] > 0
            continue
        if t == G.GmlxJobQuality:
            q = O.tet_quality64(tet, crd)
            tol_min = 1e-5
        else:
            deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
            q = O.tet_quality64(tet, O.smooth32(tet, ref, deg, ball, crd, int(p), 0.2))
            tol_min = 1e-4
        emin = abs(r[0] - q.min()) / abs(q.min())
        emean = abs(r[1] - math.fsum(q) / len(q)) / abs(math.fsum(q) / len(q))
        log("mesh %d (%d tets), job %d: min rel err %.2e, mean rel err %.2e" % (i, len(tet), t, emin, emean))
        assert emin <= tol_min and emean <= 1e-5


def test_spmv_256cube_vs_oracle():
    # the edge Laplacian of the 256^3 cube (the 100.7M-tet mesh): 16,974,593
    # block rows, 253,036,801 3x3 blocks; y = A x against the oracle's FP64
    # product over every row
    n = 256
    crd, ref, tet = G.cube_mesh(n, 0.2, 2024, True)
    V, T = len(crd), len(tet)
    L = G.lib()
    rowptr = np.empty(V + 1, np.int32)
    nb = L.GmlxEdgeLaplacian3(V, T, G.ptr(tet), 2024, 0.01, G.ptr(rowptr), None, None, 0)
    assert nb == V + 2 * 118031104
    col = np.empty(nb, np.int32)
    val = np.empty(nb * 9, np.float32)
    assert L.GmlxEdgeLaplacian3(V, T, G.ptr(tet), 2024, 0.01, G.ptr(rowptr), G.ptr(col), G.ptr(val), nb) == nb
    del crd, ref, tet
    x = np.random.default_rng(11).normal(size=V * 3).astype(np.float32)
    with G.Gml() as g:
        m = L.GmlNewMatrix(g.h, V, nb, 3, G.ptr(val), G.ptr(col), G.ptr(rowptr), G.GmlFlt)
        vx = L.GmlNewVector(g.h, V, 3, G.ptr(x), G.GmlFlt)
        vy = L.GmlNewVector(g.h, V, 3, None, G.GmlFlt)
        assert min(m, vx, vy) > 0
        assert L.GmlMultMatVec(g.h, m, vy, vx) == 1
        y = np.empty(V * 3, np.float32)
        assert L.GmlxGetVector(g.h, vy, G.ptr(y))
    y64 = O.bsr_spmv64(rowptr, col, val, x.astype(np.float64), 3)
    err = float(np.abs(y - y64).max() / np.abs(y64).max())
    log("SpMV 256^3: max |y - y64| / max |y64| = %.2e" % err)
    assert err <= 1e-5
