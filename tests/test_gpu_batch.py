"""GPU parity of the batched mesh-quality engine (GmlxNewBatch /
GmlxBatchQuality): a batch of uneven cube meshes processed by the persistent
tile kernel from its HIP graph must give, mesh by mesh,
  - per-tet quality bit-identical to GmlxTetQuality on that mesh alone (the
    same device expression), and within 1e-5 relative of the FP64 oracle,
  - per-vertex ball minimum bit-identical to GmlxBallMin,
  - per-mesh minimum equal to the minimum of those qualities, and per-mesh
    mean within 1e-12 relative of the FP64 mean of the FP32 qualities,
for every launch shape (tile size, ball padding), the autotuned one included,
and identically on every replay of the graph, with other GPU work between
the replays."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def meshes(sizes, seed=5):
    return [G.cube_mesh(n, 0.3, seed + i) for i, n in enumerate(sizes)]


def per_mesh_reference(ms):
    """(quality, ball min) of each mesh alone, via GmlxTetQuality / GmlxBallMin."""
    out = []
    for crd, ref, tet in ms:
        with G.Gml() as g:
            g.set_mesh(crd, ref, tet)
            q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
            vq = g.solution(G.GmlVertices, 1, G.GmlFlt, "vqal")
            assert g.L.GmlxTetQuality(g.h, q) == 1
            assert g.L.GmlxBallMin(g.h, q, vq) == 1
            out.append((g.download_raw(q, np.float32, (len(tet),)), g.download_raw(vq, np.float32, (len(crd),))))
    return out


def run_batch(g, b, nm, V, T):
    mn = np.empty(nm, np.float32)
    mean = np.empty(nm, np.float64)
    assert g.L.GmlxBatchQuality(g.h, b, G.ptr(mn), G.ptr(mean)) == 1
    tq = np.empty(T, np.float32)
    vm = np.empty(V, np.float32)
    assert g.L.GmlxGetBatchField(g.h, b, G.ptr(tq), G.ptr(vm)) == 1
    return mn, mean, tq, vm


def check(ms, refs, mn, mean, tq, vm):
    vo = to = 0
    for i, ((crd, ref, tet), (q0, v0)) in enumerate(zip(ms, refs)):
        q = tq[to:to + len(tet)]
        v = vm[vo:vo + len(crd)]
        assert np.array_equal(q, q0), "mesh %d: tet quality differs from GmlxTetQuality" % i
        assert np.array_equal(v, v0), "mesh %d: ball min differs from GmlxBallMin" % i
        assert mn[i] == q0.min()
        assert mean[i] == pytest.approx(q0.astype(np.float64).mean(), rel=1e-12)
        vo += len(crd)
        to += len(tet)


SIZES = [2, 7, 3, 11, 1, 5, 9, 4, 6, 13, 2, 8]


@pytest.fixture(scope="module")
def data():
    ms = meshes(SIZES)
    return ms, per_mesh_reference(ms), G.concat_meshes(ms)


@pytest.mark.parametrize("tv,pad", [(64, 1), (256, 2), (512, 4), (1024, 1)])
def test_batch_matches_per_mesh(data, tv, pad):
    ms, refs, (nver, ntet, crd, tet) = data
    V, T = int(nver.sum()), int(ntet.sum())
    with G.Gml() as g:
        b = g.L.GmlxNewBatch(g.h, len(ms), G.ptr(nver), G.ptr(ntet), G.ptr(crd), G.ptr(tet))
        assert b > 0
        assert g.L.GmlxBatchSetShape(g.h, b, tv, pad) == 1
        info = (C.c_longlong * 12)()
        assert g.L.GmlxGetBatchInfo(g.h, b, info)
        assert list(info[:3]) == [len(ms), V, T] and info[4] == tv and info[5] == pad and info[10] == 1
        assert info[7] >= T and info[8] >= V          # every tet and vertex is in some tile
        check(ms, refs, *run_batch(g, b, len(ms), V, T))
        assert g.L.GmlxFreeBatch(g.h, b) == 1


def test_batch_tuned_graph_replays(data):
    ms, refs, (nver, ntet, crd, tet) = data
    V, T = int(nver.sum()), int(ntet.sum())
    other_crd, other_ref, other_tet = G.cube_mesh(10, 0.2, 99)
    with G.Gml() as g, G.Gml() as h2:
        b = g.L.GmlxNewBatch(g.h, len(ms), G.ptr(nver), G.ptr(ntet), G.ptr(crd), G.ptr(tet))
        assert b > 0
        tv, pad = C.c_int(), C.c_int()
        assert g.L.GmlxBatchTune(g.h, b, C.byref(tv), C.byref(pad)) == 1
        assert tv.value in (256, 512, 1024) and pad.value in (1, 2, 4)
        tab = (C.c_double * 27)()
        n = g.L.GmlxGetBatchTuning(g.h, b, tab)
        assert n == 9 and all(tab[3 * i + 2] > 0 for i in range(n))
        # eager work on another handle between the replays
        h2.set_mesh(other_crd, other_ref, other_tet)
        q2 = h2.solution(G.GmlTetrahedra, 1, G.GmlFlt, "q2")
        first = None
        for rep in range(4):
            out = run_batch(g, b, len(ms), V, T)
            check(ms, refs, *out)
            if first is None:
                first = out
            else:
                for a, c in zip(first, out):
                    assert np.array_equal(a, c)
            assert h2.L.GmlxTetQuality(h2.h, q2) == 1
            h2.L.GmlxSync(h2.h)
        cnt = C.c_longlong()
        assert g.L.GmlxGetBuiltinRunTime(g.h, G.GmlxBuiltinBatch, C.byref(cnt)) > 0 and cnt.value == 4


def test_batch_quality_vs_fp64_oracle(data):
    """FP32 mean ratio against the FP64 oracle on the same FP32 coordinates.
    The volume's FP32 cancellation gives an absolute error of order
    eps (1 + |q|^-1/2) (q = 12 (9 v^2)^(1/3) / s, dv/v ~ eps / q^1.5): the
    bound is 8 eps (1 + |q|^-1/2) for every tet, and 1e-5 relative for tets
    with |q| >= 0.05; the jitter of +-0.3h makes some tets nearly flat or
    inverted, where only the first bound is meaningful."""
    eps = 2.0 ** -24
    worst_rel = worst_scaled = 0.
    for (crd, ref, tet), (q0, _) in zip(ms_of(data), refs_of(data)):
        q64 = O.tet_quality64(tet, crd)
        ae = np.abs(q0 - q64)
        scaled = ae / (eps * (1 + 1 / np.sqrt(np.maximum(np.abs(q64), 1e-12))))
        good = np.abs(q64) >= 0.05
        rel = ae[good] / np.abs(q64[good])
        worst_scaled = max(worst_scaled, scaled.max())
        worst_rel = max(worst_rel, rel.max() if rel.size else 0.)
    print("max |q - q64| / (eps (1 + |q64|^-1/2)) = %.2f; max relative error (|q| >= 0.05) = %.2e"
          % (worst_scaled, worst_rel))
    assert worst_scaled <= 8
    assert worst_rel <= 1e-5


def ms_of(data):
    return data[0]


def refs_of(data):
    return data[1]


def test_batch_rejects_bad_input():
    crd, ref, tet = G.cube_mesh(2, 0.1, 1)
    nver = np.array([len(crd)], np.int32)
    ntet = np.array([len(tet)], np.int32)
    bad = tet.copy()
    bad[3, 2] = len(crd)                     # vertex index out of its mesh
    with G.Gml() as g:
        assert g.L.GmlxNewBatch(g.h, 1, G.ptr(nver), G.ptr(ntet), G.ptr(crd), G.ptr(bad)) == 0
        b = g.L.GmlxNewBatch(g.h, 1, G.ptr(nver), G.ptr(ntet), G.ptr(crd), G.ptr(tet))
        assert b > 0
        assert g.L.GmlxBatchSetShape(g.h, b, 100, 1) == -2
        assert g.L.GmlxBatchSetShape(g.h, b, 256, 3) == -2
        assert g.L.GmlxBatchQuality(g.h, b + 1, None, None) == -1
