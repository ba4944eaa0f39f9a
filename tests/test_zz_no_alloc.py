"""GmlxReduceVectorNoAlloc: the reduction with its partials in a file-scope
device array (no allocation at call time) gives exactly what
GmlReduceVector gives, for every operation, on two calls in a row the second
of which reduces a smaller vector."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G

pytestmark = pytest.mark.gpu


def test_no_alloc_reduce_matches_reduce_vector():
    f = G.lib().GmlxReduceVectorNoAlloc
    rng = np.random.default_rng(11)
    # a large tet field (many work-groups, capped grid) then a small vertex
    # field whose partials are far fewer than the first call left behind
    crd, ref, tet = G.cube_mesh(40, 0.2, 5)
    big = rng.normal(size=len(tet)).astype(np.float32)
    big[rng.integers(len(tet), size=1000)] = 0.0
    small = rng.normal(size=len(crd)).astype(np.float32)[:len(crd)]
    with G.Gml(0) as g:
        g.set_mesh(crd, ref, tet)
        qb = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "big")
        qs = g.solution(G.GmlVertices, 1, G.GmlFlt, "small")
        g.upload(qb, big)
        g.upload(qs, small)
        for op in range(G.GmlMaxRed):
            for idx, arr in ((qb, big), (qs, small)):
                a, b = C.c_double(), C.c_double()
                assert f(g.h, idx, op, C.byref(a)) == 1
                assert g.L.GmlReduceVector(g.h, idx, op, C.byref(b)) == 1
                # same kernels, same partial order: the same double
                assert a.value == b.value, (op, len(arr), a.value, b.value)
                if op == G.GmlMin:
                    assert a.value == float(arr.min())
                if op == G.GmlMax:
                    assert a.value == float(arr.max())
                if op == G.GmlL0:
                    assert a.value == float(np.count_nonzero(arr))
        # the error codes of GmlReduceVector
        r = C.c_double()
        assert f(g.h, 0, G.GmlMin, C.byref(r)) == -1
        assert f(g.h, qb, 7, C.byref(r)) == -3
        assert f(g.h, qb, G.GmlMin, None) == -4
