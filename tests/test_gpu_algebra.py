"""GPU parity of the block-sparse algebra: GmlNewMatrix / GmlMultMatVec with
3x3 (and 4x4, 5x5) blocks against the FP64 oracle and scipy, the vector ops
against numpy, and block-Jacobi CG in FP32 and with FP64 refinement against
an FP64 direct solve.  Tolerances: FP32 products within 1e-5 of max|y|
(FP32 accumulation over <= 15 blocks); refined CG to 1e-10 relative."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import gmlib_amd as G
import gmlib_amd.workloads as W
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def laplacian(n, seed=1):
    crd, ref, tet = G.cube_mesh(n, 0.2, seed)
    e = W.cube_edges(tet)
    rp, col, val = W.edge_laplacian3(len(crd), e, seed)
    return len(crd), rp, col, val


def new_vec(g, n, b, vals=None, flt=G.GmlFlt):
    dt = np.float32 if flt == G.GmlFlt else np.float64
    a = np.ascontiguousarray(vals, dt) if vals is not None else None
    i = g.L.GmlNewVector(g.h, n, b, G.ptr(a) if a is not None else None, flt)
    assert i > 0
    return i


def get_vec(g, i, n, b, dt=np.float32):
    out = np.empty(n * b, dt)
    assert g.L.GmlxGetVector(g.h, i, G.ptr(out))
    return out


@pytest.mark.parametrize("b", [3, 4, 5])
def test_spmv_blocks_vs_oracle(b):
    nv, rp, col, val3 = laplacian(6)
    rng = np.random.default_rng(b)
    val = rng.normal(size=(len(col), b, b)).astype(np.float32)
    x = rng.normal(size=nv * b).astype(np.float32)
    with G.Gml() as g:
        m = g.L.GmlNewMatrix(g.h, nv, len(col), b, G.ptr(val), G.ptr(col), G.ptr(rp), G.GmlFlt)
        assert m > 0
        vx = new_vec(g, nv, b, x)
        vy = new_vec(g, nv, b)
        assert g.L.GmlMultMatVec(g.h, m, vy, vx) == 1      # vec1 = A . vec2
        y = get_vec(g, vy, nv, b)
        assert g.L.GmlGetMemoryAccess(g.h) > 0
    y64 = O.bsr_spmv64(rp, col, val, x.astype(np.float64), b)
    M = sp.bsr_matrix((val.astype(np.float64), col, rp), shape=(nv * b, nv * b))
    assert np.allclose(y64, M @ x.astype(np.float64), rtol=1e-12, atol=1e-12)
    err = np.abs(y - y64).max() / np.abs(y64).max()
    print("b=%d SpMV max err / max|y| vs FP64 oracle: %.3e" % (b, err))
    assert err <= 1e-5


def test_spmv_fp64_values_and_errors():
    nv, rp, col, val = laplacian(4)
    x = np.random.default_rng(0).normal(size=nv * 3)
    val64 = val.astype(np.float64)         # must outlive the call: ptr() does not hold a reference
    with G.Gml() as g:
        m = g.L.GmlNewMatrix(g.h, nv, len(col), 3, G.ptr(val64), G.ptr(col), G.ptr(rp), G.GmlDbl)
        vx = new_vec(g, nv, 3, x, G.GmlDbl)
        vy = new_vec(g, nv, 3, None, G.GmlDbl)
        assert g.L.GmlMultMatVec(g.h, m, vy, vx) == 1
        y = get_vec(g, vy, nv, 3, np.float64)
        small = new_vec(g, nv - 1, 3, None, G.GmlDbl)
        assert g.L.GmlMultMatVec(g.h, m, small, vx) == -4
        assert g.L.GmlMultMatVec(g.h, 9, vy, vx) == -1
    y64 = O.bsr_spmv64(rp, col, val, x, 3)
    assert np.allclose(y, y64, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [1000, 1001])
def test_vector_ops(n):
    b = 3
    rng = np.random.default_rng(1)
    u, v, w = (rng.normal(size=n * b).astype(np.float32) for _ in range(3))
    D = rng.normal(size=(n, b, b)).astype(np.float32)
    with G.Gml() as g:
        L, h = g.L, g.h
        iu, iv, iw, iy = new_vec(g, n, b, u), new_vec(g, n, b, v), new_vec(g, n, b, w), new_vec(g, n, b)
        idg = new_vec(g, n, b * b, D.reshape(-1))
        assert L.GmlAddVec3(h, iu, iv, iw, iy) == 1
        assert np.allclose(get_vec(g, iy, n, b), u + v + w, rtol=1e-6, atol=1e-6)
        assert L.GmlMultDiagMatVec(h, idg, iu, iy) == 1
        assert np.allclose(get_vec(g, iy, n, b), np.einsum("rij,rj->ri", D, u.reshape(n, b)).ravel(), rtol=1e-5, atol=1e-5)
        assert L.GmlMultDiagMatVec(h, iu, iu, iy) == -4
        f = C.c_double(-2.5)
        assert L.GmlScaleVec(h, iu, C.byref(f)) == 1
        assert np.array_equal(get_vec(g, iu, n, b), (u * np.float32(-2.5)))
        L.GmlNewMeshData(h, G.GmlVertices, n)
        red = g.solution(G.GmlVertices, 1, G.GmlFlt, "reduction")
        nrm = C.c_double()
        assert L.GmlNormVec(h, iv, red, C.byref(nrm)) == 1
        want = np.sqrt(np.sum(np.linalg.norm(v.reshape(n, b).astype(np.float64), axis=1)))
        assert nrm.value == pytest.approx(want, rel=1e-6)
        d = C.c_double()
        assert L.GmlxDotVec(h, iv, iw, C.byref(d)) == 1
        assert d.value == pytest.approx(float(np.dot(v.astype(np.float64), w.astype(np.float64))), rel=1e-12)


@pytest.mark.parametrize("n,b", [(1, 1), (7, 3), (513, 3), (1000, 2)])
def test_vector_ops_fp64_tails(n, b):
    """16-byte vector kernels: sizes whose entry count is odd or not a
    multiple of the vector width take the scalar tail."""
    rng = np.random.default_rng(n)
    u, v, w = (rng.normal(size=n * b) for _ in range(3))
    with G.Gml() as g:
        L, h = g.L, g.h
        iu, iv, iw = (new_vec(g, n, b, a, G.GmlDbl) for a in (u, v, w))
        iy = new_vec(g, n, b, None, G.GmlDbl)
        assert L.GmlAddVec3(h, iu, iv, iw, iy) == 1
        assert np.array_equal(get_vec(g, iy, n, b, np.float64), (u + v) + w)
        f = C.c_double(0.75)
        assert L.GmlScaleVec(h, iu, C.byref(f)) == 1
        assert np.array_equal(get_vec(g, iu, n, b, np.float64), u * 0.75)
        fu, fv, fw = (a.astype(np.float32) for a in (u, v, w))
        ju, jv, jw = (new_vec(g, n, b, a) for a in (fu, fv, fw))
        jy = new_vec(g, n, b)
        assert L.GmlAddVec3(h, ju, jv, jw, jy) == 1
        assert np.array_equal(get_vec(g, jy, n, b), (fu + fv) + fw)
        assert L.GmlScaleVec(h, ju, C.byref(f)) == 1
        assert np.array_equal(get_vec(g, ju, n, b), fu * np.float32(0.75))


@pytest.mark.parametrize("refine,tol", [(0, 1e-6), (1, 1e-10)])
def test_block_cg(refine, tol):
    nv, rp, col, val = laplacian(10, 3)
    rhs = np.random.default_rng(7).normal(size=nv * 3).astype(np.float32)
    with G.Gml() as g:
        m = g.L.GmlNewMatrix(g.h, nv, len(col), 3, G.ptr(val), G.ptr(col), G.ptr(rp), G.GmlFlt)
        vb = new_vec(g, nv, 3, rhs)
        vx = new_vec(g, nv, 3, None, G.GmlDbl)
        rel, it = C.c_double(), C.c_int()
        assert g.L.GmlxSolveCG(g.h, m, vb, vx, tol, 20000, refine, C.byref(rel), C.byref(it)) == 1
        x = get_vec(g, vx, nv, 3, np.float64)
    A = sp.bsr_matrix((val.astype(np.float64), col, rp), shape=(nv * 3, nv * 3)).tocsr()
    b64 = rhs.astype(np.float64)
    r = np.linalg.norm(b64 - A @ x) / np.linalg.norm(b64)
    xs = spl.spsolve(A.tocsc(), b64)
    fe = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    print("CG refine=%d: %d iterations, device relres %.3e, host FP64 relres %.3e, |x - x_direct|/|x| %.3e"
          % (refine, it.value, rel.value, r, fe))
    assert abs(r - rel.value) <= 1e-3 * max(r, 1e-16) + 1e-15
    if refine:
        assert r <= tol and fe <= 1e-7
    else:
        # FP32 CG stops on its recurrence residual; the true residual of an
        # FP32 iterate bottoms out near cond(A) * 2^-24
        assert r <= 1e-4 and fe <= 1e-5


def permuted(rp, col, val, perm):
    """The block matrix P A P^T: new row i is old row perm[i], its columns
    renumbered by the inverse permutation (block order within a row kept)."""
    n = len(rp) - 1
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    deg = np.diff(rp)[perm]
    nrp = np.zeros(n + 1, np.int32)
    nrp[1:] = np.cumsum(deg)
    idx = np.repeat(rp[perm] - nrp[:-1], deg) + np.arange(nrp[-1])
    return nrp, inv[col[idx]].astype(np.int32), val[idx]


@pytest.mark.parametrize("b,flt", [(3, G.GmlFlt), (2, G.GmlDbl)])
def test_spmv_far_columns_and_ragged_rows(b, flt):
    """Rows whose columns lie beyond +-32767 of the row take the 32-bit column
    path, the near ones the 16-bit offsets; degrees differ inside a slab
    (so the slab's 16-bit slot count is its smallest near count), some rows
    are empty, near and far columns alternate in the caller's order."""
    n = 70001
    rng = np.random.default_rng(11 + b)
    deg = rng.integers(0, 9, size=n)
    deg[::97] = 0
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(deg)
    rows = np.repeat(np.arange(n), deg)
    near = rng.random(len(rows)) < 0.7
    # far columns about half the matrix away (|col - row| >= 34000)
    col = np.where(near, np.clip(rows + rng.integers(-32767, 32768, len(rows)), 0, n - 1),
                   (rows + n // 2 + rng.integers(-1000, 1000, len(rows))) % n).astype(np.int32)
    dt = np.float32 if flt == G.GmlFlt else np.float64
    val = rng.normal(size=(len(col), b, b)).astype(dt)
    x = rng.normal(size=n * b).astype(dt)
    with G.Gml() as g:
        m = g.L.GmlNewMatrix(g.h, n, len(col), b, G.ptr(val), G.ptr(col), G.ptr(rp), flt)
        assert m > 0
        vx = new_vec(g, n, b, x, flt)
        vy = new_vec(g, n, b, None, flt)
        assert g.L.GmlMultMatVec(g.h, m, vy, vx) == 1
        y = get_vec(g, vy, n, b, dt)
    far = np.abs(col.astype(np.int64) - rows) > 32767
    assert far.mean() > 0.2
    # FP64 products: the oracle's (FP32 block values) for FP32 data, scipy's
    # BSR product for FP64 data
    if flt == G.GmlFlt:
        y64 = O.bsr_spmv64(rp, col, val, x.astype(np.float64), b)
    else:
        y64 = sp.bsr_matrix((val, col, rp), shape=(n * b, n * b)) @ x
    err = np.abs(y - y64).max() / np.abs(y64).max()
    print("b=%d %s: %.1f%% far columns, max err / max|y| vs FP64 oracle %.3e"
          % (b, dt.__name__, 100 * far.mean(), err))
    assert err <= (1e-5 if flt == G.GmlFlt else 1e-12)
    assert np.all(y[np.repeat(deg == 0, b)] == 0)


def test_block_cg_far_columns():
    """Refined CG on a randomly renumbered edge Laplacian (68,921 block rows):
    a quarter of the couplings lie beyond +-32767 rows, so both column paths
    of the CG product run; the FP64 residual is recomputed on the host."""
    nv, rp, col, val = laplacian(40, 5)
    perm = np.random.default_rng(3).permutation(nv)
    rp, col, val = permuted(rp, col, val, perm)
    rows = np.repeat(np.arange(nv), np.diff(rp))
    assert (np.abs(col.astype(np.int64) - rows) > 32767).mean() > 0.15
    rhs = np.random.default_rng(8).normal(size=nv * 3).astype(np.float32)
    with G.Gml() as g:
        m = g.L.GmlNewMatrix(g.h, nv, len(col), 3, G.ptr(val), G.ptr(col), G.ptr(rp), G.GmlFlt)
        vb = new_vec(g, nv, 3, rhs)
        vx = new_vec(g, nv, 3, None, G.GmlDbl)
        rel, it = C.c_double(), C.c_int()
        assert g.L.GmlxSolveCG(g.h, m, vb, vx, 1e-10, 20000, 1, C.byref(rel), C.byref(it)) == 1
        x = get_vec(g, vx, nv, 3, np.float64)
    b64 = rhs.astype(np.float64)
    r = np.linalg.norm(b64 - O.bsr_spmv64(rp, col, val, x, 3)) / np.linalg.norm(b64)
    print("permuted CG: %d iterations, device relres %.3e, host FP64 relres %.3e" % (it.value, rel.value, r))
    assert r <= 1e-10
