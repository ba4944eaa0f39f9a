"""The GpuBandwidth access patterns (examples/GpuBandwidth: direct, indirect
tet -> vertex, ragged vertex -> tet ball) in the oracle, pinned by exact
integer arithmetic (small integer coordinates make every FP32 sum exact,
so a dropped, repeated or misindexed term changes the result) and by a
brute-force ball built from the tet table's positions independently of the
ball builder; and the kernel bodies compiled for gfx950 on a host-only
handle.  The GPU runs are in test_gpu_bandwidth.py."""
import os

import numpy as np

import gmlib_amd as G
from tests import oracle_lib as O

KDIR = os.path.join(os.path.dirname(G.__file__), "kernels")


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


def random_numbering(crd, ref, tet, seed):
    """the same mesh with vertices and tets in a random order"""
    rng = np.random.default_rng(seed)
    pv = rng.permutation(len(crd))              # new vertex i = old pv[i]
    inv = np.empty_like(pv)
    inv[pv] = np.arange(len(pv))
    pt = rng.permutation(len(tet))
    return (np.ascontiguousarray(crd[pv]), np.ascontiguousarray(ref[pv]),
            np.ascontiguousarray(inv[tet[pt]].astype(np.int32)))


def integer_mesh(n, seed):
    crd, ref, tet = G.cube_mesh(n, 0.2, seed)
    rng = np.random.default_rng(seed)
    icrd = np.zeros_like(crd)
    icrd[:, :3] = rng.integers(-1000, 1000, size=(len(crd), 3)).astype(np.float32)
    return random_numbering(icrd, ref, tet, seed)


def test_direct_and_indirect_exact():
    crd, ref, tet = integer_mesh(6, 3)
    d = O.bw_direct_tets(tet)
    assert np.array_equal(d, tet.astype(np.float32))
    ind = O.bw_indirect_tets(tet, crd)
    want = crd[tet].astype(np.int64).sum(axis=1)               # exact integer sums
    assert np.array_equal(ind.astype(np.int64), want)
    assert (ind[:, 3] == 0).all()


def test_ragged_matches_brute_force_ball():
    crd, ref, tet = integer_mesh(6, 4)
    V, T = len(crd), len(tet)
    tdat = np.zeros((T, 4), np.float32)
    tdat[:, :3] = np.random.default_rng(5).integers(-500, 500, size=(T, 3))
    tdat[:, 3] = np.arange(T) % 7
    deg, ball, md = O.vertex_ball(tet, V, 32)
    assert md <= 32
    got = O.bw_vertices_ball(deg, ball, tdat)
    want = np.zeros((V, 4), np.int64)
    np.add.at(want, tet.reshape(-1), np.repeat(tdat.astype(np.int64), 4, axis=0))   # one term per tet position
    assert np.array_equal(got.astype(np.int64), want)


def test_ragged_fp32_order_close_to_fp64():
    crd, ref, tet = G.cube_mesh(5, 0.3, 9)
    deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
    tdat = O.bw_indirect_tets(tet, crd)
    got = O.bw_vertices_ball(deg, ball, tdat)
    want = np.zeros((len(crd), 4))
    np.add.at(want, tet.reshape(-1), np.repeat(tdat.astype(np.float64), 4, axis=0))
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


def test_bandwidth_bodies_compile():
    crd, ref, tet = G.cube_mesh(3, 0.2, 1)
    R, W = G.GmlReadMode, G.GmlWriteMode
    with G.Gml(host_only=True) as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        vdat = g.solution(G.GmlVertices, 1, G.GmlFlt4, "VerDat")
        tdat = g.solution(G.GmlTetrahedra, 1, G.GmlFlt4, "TetDat")
        for name, typ, args in (("bw_direct_vertices", G.GmlVertices, [(vi, R, 0), (vdat, W, 0)]),
                                ("bw_direct_tets", G.GmlTetrahedra, [(ti, R, 0), (tdat, W, 0)]),
                                ("bw_indirect_tets", G.GmlTetrahedra, [(vi, R, 0), (tdat, W, 0)]),
                                ("bw_vertices_ball", G.GmlVertices, [(vdat, W, 0), (tdat, R, 0)])):
            assert g.compile(body(name), name, typ, args) > 0, name
