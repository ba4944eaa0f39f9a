"""The oracle pinned against closed forms, independent formulas, brute force
and library routines (the reference ships no golden vectors for this path).

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign or index, a transposed operand) fails it.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from tests import oracle_lib as O

REG = [np.array([1., 1., 1.]), np.array([-1., 1., -1.]), np.array([1., -1., -1.]), np.array([-1., -1., 1.])]  # positively oriented
KUHN = [np.array([0., 0., 0.]), np.array([1., 0., 0.]), np.array([1., 1., 0.]), np.array([1., 1., 1.])]


def meanratio_jacobian(p):
    """Mean ratio from the Jacobian map to the regular tet (Liu & Joe; Knupp):
    q = 3 det(S)^(2/3) / ||S||_F^2, S = J W^{-1}."""
    J = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1)
    W = np.array([[1, .5, .5], [0, math.sqrt(3) / 2, math.sqrt(3) / 6], [0, 0, math.sqrt(2. / 3.)]])
    S = J @ np.linalg.inv(W)
    d = np.linalg.det(S)
    return math.copysign(3 * abs(d) ** (2. / 3.) / np.sum(S * S), d)


def rot(ax, ang):
    ax = ax / np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def test_meanratio_closed_forms():
    assert O.meanratio64(*REG) == pytest.approx(1.0, abs=1e-14)
    assert O.meanratio64(*KUHN) == pytest.approx(1.2 * 0.5 ** (2. / 3.), abs=1e-14)
    assert O.meanratio64(*KUHN) == pytest.approx(0.7559526, abs=1e-7)
    # inverted tet: sign of the volume
    assert O.meanratio64(KUHN[0], KUHN[2], KUHN[1], KUHN[3]) == pytest.approx(-0.7559526, abs=1e-7)


def test_meanratio_vs_jacobian_form_random():
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = [rng.normal(size=3) for _ in range(4)]
        assert O.meanratio64(*p) == pytest.approx(meanratio_jacobian(p), rel=1e-10, abs=1e-12)


def test_meanratio_invariances():
    rng = np.random.default_rng(4)
    p = [rng.normal(size=3) for _ in range(4)]
    q = O.meanratio64(*p)
    R = rot(np.array([.3, -1., 2.]), 1.1)
    t = np.array([5., -3., 7.])
    assert O.meanratio64(*[3.7 * R @ x + t for x in p]) == pytest.approx(q, rel=1e-12)
    # even permutation of the vertices keeps it, an odd one flips the sign
    assert O.meanratio64(p[1], p[2], p[0], p[3]) == pytest.approx(q, rel=1e-12)
    assert O.meanratio64(p[1], p[0], p[2], p[3]) == pytest.approx(-q, rel=1e-12)


def test_meanratio_f32_operation_order():
    rng = np.random.default_rng(5)
    for _ in range(200):
        p = [rng.uniform(0, 1, size=3) for _ in range(4)]
        f = [O.F4(*[float(v) for v in x], 0.0) for x in p]
        q32 = O.lib().orc_meanratio_f32(*f)
        q64 = O.meanratio64(*[np.float32(x).astype(np.float64) for x in p])
        assert abs(q32 - q64) <= 2e-6 * max(1.0, abs(q64))


def test_reference_quality_formulas_closed_forms():
    L = O.lib()
    P = lambda x: O.p(O.d3(x))
    # examples/MeshQuality/compute_quality.cl: only two faces are summed and
    # 3*sqrt(6) makes the regular tet come out at 1 (0.999996 with the
    # truncated constants)
    r = L.orc_ref_compute_quality_f64(*[P(x) for x in REG])
    assert r == pytest.approx(0.999996, abs=2e-6)
    k = L.orc_ref_compute_quality_f64(*[P(x) for x in KUHN])
    # Kuhn: V=1/6, lmax^2=3, faces (0,1,2) and (0,1,3): areas 1/2 and sqrt(2)/2
    assert k == pytest.approx(7.348469 * (0.166666 * 1.0) / ((0.5 + math.sqrt(2) / 2) * math.sqrt(3)), rel=1e-12)
    # toolkit CalTetQal keeps the repeated (b,c) edge: regular tet -> 6
    t = L.orc_caltetqal_f64(*[P(x) for x in REG])
    assert t == pytest.approx(5.99999, abs=2e-5)
    h = 1 + 1 + math.sqrt(3) + math.sqrt(2) + math.sqrt(2) + 1  # (a,b)(a,c)(a,d)(b,c)(b,c)(c,d)
    # Kuhn with (a,b,c,d) = (0,0,0),(1,0,0),(1,1,0),(1,1,1)
    h = 1.0 + math.sqrt(2) + math.sqrt(3) + 1.0 + 1.0 + 1.0
    s = 1.0 + math.sqrt(2) + 1.0 + math.sqrt(2)   # |cross| of faces abc, cad, bcd, dab
    assert L.orc_caltetqal_f64(*[P(x) for x in KUHN]) == pytest.approx(176.363 * 1.0 / (h * s), rel=1e-12)
    e = [np.array([0., 0., 0.]), np.array([1., 0., 0.]), np.array([.5, math.sqrt(3) / 2, 0.])]
    assert L.orc_caltriqal_f64(*[P(x) for x in e]) == pytest.approx(1.0, abs=2e-5)


def brute_ball(tet, nver):
    rows = [[] for _ in range(nver)]
    for e in range(len(tet) - 1, -1, -1):
        for j in range(tet.shape[1]):
            rows[tet[e, j]].append((e << 4) | j)
    return rows


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_vertex_ball_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    nver = 30
    tet = np.array([rng.choice(nver, 4, replace=False) for _ in range(120)], np.int32)
    deg, ball, md = O.vertex_ball(tet, nver, 64)
    rows = brute_ball(tet, nver)
    assert md == max(len(r) for r in rows)
    for v in range(nver):
        assert deg[v] == len(rows[v])
        assert list(ball[v, :deg[v]]) == rows[v]
        assert not ball[v, deg[v]:].any()
    # descending element order, local index recovers the vertex
    for v in range(nver):
        e = ball[v, :deg[v]] >> 4
        assert np.all(np.diff(e) < 0)
        assert np.all(tet[e, ball[v, :deg[v]] & 15] == v)


def test_vertex_ball_sum_of_degrees():
    rng = np.random.default_rng(7)
    tet = np.array([rng.choice(50, 4, replace=False) for _ in range(300)], np.int32)
    deg, _, _ = O.vertex_ball(tet, 50, 128)
    assert deg.sum() == 4 * len(tet)


def test_get_cnt_vec_table():
    # gmlib3.c:2945-2964: p = ceil(log2 n) capped at 7
    want = {1: (1, 1), 2: (1, 2), 3: (1, 4), 4: (1, 4), 5: (1, 8), 8: (1, 8), 9: (1, 16), 16: (1, 16),
            17: (2, 16), 32: (2, 16), 33: (4, 16), 64: (4, 16), 65: (8, 16), 128: (8, 16), 500: (8, 16)}
    import ctypes as C
    for n, (c, v) in want.items():
        a, b = C.c_int(0), C.c_int(0)
        O.lib().orc_get_cnt_vec(n, C.byref(a), C.byref(b))
        assert (a.value, b.value) == (c, v), n


def numbering(first_nodes):
    """orc_evaluate_numbering over elements whose first node is given (the
    reference reads only the first node of each element line: NmbItm = 1,
    gmlib3.c:589 and :4423-4426)."""
    e = np.zeros((len(first_nodes), 4), np.int32)
    e[:, 0] = first_nodes
    e[:, 1:] = 987654                       # never read
    return O.lib().orc_evaluate_numbering(len(e), 4, O.p(e))


def test_evaluate_numbering_closed_forms():
    # gmlib3.c:4409-4442: a 16-entry direct-mapped tag table indexed by bits
    # 4..7 of the vertex id, tag = id >> 10, tags start at 0
    n = 4096
    # one vertex id below 1024: its tag is the initial 0, every access hits
    assert numbering(np.full(n, 700)) == 100.0
    # one id above 1024: one miss, then hits
    assert numbering(np.full(n, 5000)) == pytest.approx(100.0 * (n - 1) / n, rel=1e-6)
    # ids 1024 apart share a slot with a new tag each time: every access misses
    assert numbering(1024 * np.arange(1, n + 1)) == 0.0
    # consecutive ids 0..n-1: the first 1024 hit (tag 0); in each later block
    # of 1024 the first access to each of the 16 slots misses
    miss = 16 * (n // 1024 - 1)
    assert numbering(np.arange(n)) == pytest.approx(100.0 * (n - miss) / n, rel=1e-6)
    # ids 16 apart walk the 16 slots: 64 ids per slot per block of 1024 ids,
    # the first of each slot in each block >= 1 misses
    ids = 16 * np.arange(n)
    blocks = (ids[-1] >> 10) + 1
    assert numbering(ids) == pytest.approx(100.0 * (n - 16 * (blocks - 1)) / n, rel=1e-6)
    # two ids in different blocks whose bits 4..7 differ (slots 0 and 1) but
    # whose bits 5..8 agree: alternating them hits after the first two misses
    ab = np.tile(np.array([1024, 2048 + 16]), n // 2)
    assert numbering(ab) == pytest.approx(100.0 * (n - 2) / n, rel=1e-6)
    # the same two ids in one slot (bits 4..7 equal) evict each other
    assert numbering(np.tile(np.array([1024, 2048]), n // 2)) == 0.0
    # only the first node counts: identical first nodes with varied others
    e = np.random.default_rng(2).integers(0, 10 ** 6, size=(n, 4)).astype(np.int32)
    e[:, 0] = 300
    assert O.lib().orc_evaluate_numbering(n, 4, O.p(e)) == 100.0


def test_evaluate_numbering_shares_tags_across_kinds():
    # GmlEvaluateNumbering runs one tag table over every element kind in turn
    # (gmlib3.c:4416-4438): edges whose first node is 5000 leave the tag for
    # the tets that follow, so only the very first access misses
    import ctypes as C
    import gmlib_amd as G
    ne, nt = 50, 70
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi = L.GmlNewMeshData(h, G.GmlVertices, 6000)
        ei = L.GmlNewMeshData(h, G.GmlEdges, ne)
        ti = L.GmlNewMeshData(h, G.GmlTetrahedra, nt)
        assert vi and ei and ti
        edg = np.tile(np.array([[5000, 17, 0]], np.int32), (ne, 1))
        tet = np.tile(np.array([[5000, 1, 2, 3]], np.int32), (nt, 1))
        r = np.zeros(max(ne, nt), np.int32)
        assert L.GmlSetDataBlock(h, G.GmlEdges, 0, ne - 1, G.ptr(edg), C.c_void_p(edg.ctypes.data + (ne - 1) * 8),
                                 G.ptr(r), C.c_void_p(r.ctypes.data + (ne - 1) * 4))
        assert L.GmlSetDataBlock(h, G.GmlTetrahedra, 0, nt - 1, G.ptr(tet), C.c_void_p(tet.ctypes.data + (nt - 1) * 16),
                                 G.ptr(r), C.c_void_p(r.ctypes.data + (nt - 1) * 4))
        got = L.GmlEvaluateNumbering(h)
    assert got == pytest.approx(100.0 * (ne + nt - 1) / (ne + nt), rel=1e-6)
    # the oracle over the concatenated first nodes gives the same figure
    assert numbering(np.concatenate([edg[:, 0], tet[:, 0]])) == pytest.approx(got, rel=1e-6)


@pytest.mark.parametrize("n", [0, 1, 3, 4, 1027])
def test_reductions_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n).astype(np.float32)
    if n > 3:
        x[::5] = 0.
    if n:
        assert O.reduce(0, x) == float(np.min(x))
        assert O.reduce(1, x) == float(np.max(x))
        assert O.reduce(6, x) == float(np.max(np.abs(x)))
    else:
        assert O.reduce(0, x) == float(np.finfo(np.float32).max)
    x64 = x.astype(np.float64)
    assert O.reduce(2, x) == pytest.approx(math.fsum(x64), abs=1e-12)
    assert O.reduce(3, x) == float(np.count_nonzero(x))
    assert O.reduce(4, x) == pytest.approx(math.fsum(np.abs(x64)), abs=1e-12)
    assert O.reduce(5, x) == pytest.approx(math.fsum(x64 * x64), abs=1e-12)


def test_ball_min_brute_force():
    rng = np.random.default_rng(2)
    tet = np.array([rng.choice(40, 4, replace=False) for _ in range(150)], np.int32)
    deg, ball, _ = O.vertex_ball(tet, 41, 64)     # vertex 40 has an empty ball
    q = rng.normal(size=len(tet)).astype(np.float32)
    got = O.ball_min(deg, ball, q)
    for v in range(41):
        es = [e for e in range(len(tet)) if v in tet[e]]
        assert got[v] == (min(q[es]) if es else 0.0)


def small_mesh(n, jitter):
    import gmlib_amd as G
    return G.cube_mesh(n, jitter, 11)


def test_smoothing_fixed_point_of_uniform_grid():
    # barycentre smoothing: every interior ball of the unjittered Kuhn grid is
    # centrally symmetric, so the mean barycentre is the vertex itself
    crd, ref, tet = small_mesh(4, 0.0)
    deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
    c64 = O.smooth64(tet, ref, deg, ball, crd, 3, 0.7)
    assert np.allclose(c64, crd[:, :3], atol=1e-14)
    c32 = O.smooth32(tet, ref, deg, ball, crd, 3, 0.7)
    assert np.allclose(c32, crd, atol=1e-6)


def test_smoothing_f32_tracks_f64_and_moves_interior_only():
    crd, ref, tet = small_mesh(5, 0.2)
    deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
    c64 = O.smooth64(tet, ref, deg, ball, crd, 10, 0.2)
    c32 = O.smooth32(tet, ref, deg, ball, crd, 10, 0.2)
    assert np.abs(c32[:, :3] - c64).max() < 2e-6
    b = ref != 0
    assert np.array_equal(c32[b], crd[b])
    assert np.abs(c64[~b] - crd[~b, :3]).max() > 1e-3
    # smoothing makes the mesh better: the worst mean ratio goes up
    q0 = O.tet_quality64(tet, crd).min()
    c = np.zeros_like(crd)
    c[:, :3] = c64
    assert O.tet_quality64(tet, c.astype(np.float32)).min() > q0


def test_smoothing_hand_computed_vertex():
    # one sweep, one vertex, recomputed here from the definition
    crd, ref, tet = small_mesh(3, 0.2)
    deg, ball, _ = O.vertex_ball(tet, len(crd), 32)
    v = int(np.flatnonzero(ref == 0)[0])
    es = [e for e in range(len(tet)) if v in tet[e]]
    bary = crd[tet[es], :3].astype(np.float64).mean(axis=1)
    want = crd[v, :3] + (bary.mean(axis=0) - crd[v, :3]) * 0.3
    got = O.smooth64(tet, ref, deg, ball, crd, 1, 0.3)[v]
    assert np.allclose(got, want, atol=1e-14)


def test_reference_nodesmooth_ideal_apex_of_regular_tet():
    # examples/NodeSmoothing/scatter.cl puts the apex at height sqrt(2/3) *
    # edge along cross(edges) of the face; for a positively oriented regular
    # tet that normal points away from the opposite vertex, so the "ideal"
    # apex is the mirror image of the opposite vertex through the face
    p = np.array(REG, np.float64) + 5.0
    tet = np.array([[0, 1, 2, 3]], np.int32)
    deg, ball, _ = O.vertex_ball(tet, 4, 32)
    _, opt, _ = O.ref_nodesmooth64(tet, deg, ball, np.hstack([p, np.zeros((4, 1))]), 1)
    for f in range(4):
        c = np.delete(p, f, axis=0).mean(axis=0)
        assert np.allclose(opt[0, f], 2 * c - p[f], atol=2e-5)


def test_reference_nodesmooth_scatter_closed_form_corner_tet():
    # corner tet (0, e1, e2, e3), positively oriented.  Face opposite vertex 0
    # is equilateral with side sqrt(2): centroid (1/3,1/3,1/3), outward normal
    # (1,1,1)/sqrt(3), height 0.816496 * sqrt(2) -> ideal position ~(1,1,1).
    # Face opposite vertex 1 is a right isoceles triangle with legs 1: the
    # height uses its longest edge, the hypotenuse sqrt(2), along -x.
    p = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]]) + 3.0
    tet = np.array([[0, 1, 2, 3]], np.int32)
    deg, ball, _ = O.vertex_ball(tet, 4, 32)
    _, opt, _ = O.ref_nodesmooth64(tet, deg, ball, np.hstack([p, np.zeros((4, 1))]), 1)
    hgt = 0.816496 * math.sqrt(2.0)
    assert np.allclose(opt[0, 0], 3.0 + 1.0 / 3.0 + hgt / math.sqrt(3.0), atol=1e-12)
    assert np.allclose(opt[0, 1], [3.0 - hgt, 3.0 + 1.0 / 3.0, 3.0 + 1.0 / 3.0], atol=1e-12)
    assert np.allclose(opt[0, 2], [3.0 + 1.0 / 3.0, 3.0 - hgt, 3.0 + 1.0 / 3.0], atol=1e-12)
    assert np.allclose(opt[0, 3], [3.0 + 1.0 / 3.0, 3.0 + 1.0 / 3.0, 3.0 - hgt], atol=1e-12)


def test_reference_nodesmooth_scatter_geometry_random_tets():
    # for any positively oriented tet, each ideal position lies on the line
    # through the centroid of the face opposite vertex f, perpendicular to
    # that face, on the side away from vertex f, at 0.816496 times the
    # longest of the face's three edges
    rng = np.random.default_rng(8)
    pts, tets = [], []
    while len(tets) < 60:
        q = rng.normal(size=(4, 3))
        if np.dot(np.cross(q[1] - q[0], q[2] - q[0]), q[3] - q[0]) < 0.05:
            continue
        tets.append(np.arange(4) + 4 * len(tets))
        pts.append(q)
    P = np.vstack(pts)
    tet = np.array(tets, np.int32)
    deg, ball, _ = O.vertex_ball(tet, len(P), 32)
    _, opt, _ = O.ref_nodesmooth64(tet, deg, ball, np.hstack([P, np.zeros((len(P), 1))]), 1)
    for t in range(len(tet)):
        q = P[tet[t]]
        for f in range(4):
            face = np.delete(q, f, axis=0)
            c = face.mean(axis=0)
            longest = max(np.linalg.norm(face[i] - face[j]) for i, j in ((0, 1), (0, 2), (1, 2)))
            nrm = np.cross(face[1] - face[0], face[2] - face[0])
            nrm /= np.linalg.norm(nrm)
            if np.dot(nrm, q[f] - c) > 0:
                nrm = -nrm
            assert np.allclose(opt[t, f], c + 0.816496 * longest * nrm, atol=1e-10), (t, f)


def test_reference_nodesmooth_gather_closed_form_bipyramid():
    # two regular tets sharing face (a, b, c), apexes d and e = mirror of d.
    # In a regular tet the ideal position of a vertex is its mirror image
    # through the opposite face (up to 0.816496 / sqrt(2/3)).  Gather: the
    # ball mean of the ideal positions chosen by each entry's local index,
    # moved 0.2 of the way, residual |old - new|^2 = 0.04 |mean - old|^2.
    k = 0.816496 / math.sqrt(2.0 / 3.0)
    a, b, c, d = (np.array(x) for x in REG)
    ctr = (a + b + c) / 3.0
    e = 2 * ctr - d
    P = np.array([a, b, c, d, e]) + 2.0
    tet = np.array([[0, 1, 2, 3], [0, 2, 1, 4]], np.int32)      # both positively oriented
    deg, ball, _ = O.vertex_ball(tet, 5, 32)
    new, opt, res = O.ref_nodesmooth64(tet, deg, ball, np.hstack([P, np.zeros((5, 1))]), 1)

    def ideal(t, j):
        """mirror of vertex j of tet t through its opposite face, scaled by k"""
        q = P[tet[t]]
        f = np.delete(q, j, axis=0).mean(axis=0)
        return f + k * (f - q[j])

    for v in range(5):
        entries = [(t, list(tet[t]).index(v)) for t in range(2) if v in tet[t]]
        mean = np.mean([ideal(t, j) for t, j in entries], axis=0)
        want = P[v] + 0.2 * (mean - P[v])
        assert np.allclose(new[v], want, atol=1e-12), v
        assert res[v] == pytest.approx(0.04 * np.sum((mean - P[v]) ** 2), rel=1e-10)
    # the apexes: each moves 0.2 of the way to (k-scaled) the other apex
    assert np.allclose(new[3], P[3] + 0.2 * (1 + k) / 2 * (P[4] - P[3]), atol=1e-12)
    # the height of a regular tet of edge s is sqrt(2/3) s: residual of an apex
    s = np.linalg.norm(a - b)
    assert res[3] == pytest.approx(0.04 * ((1 + k) * math.sqrt(2.0 / 3.0) * s) ** 2, rel=1e-10)


def test_face_pull_closed_forms():
    # one regular tet: each vertex moves a quarter of the way to the centroid
    # of its opposite face, i.e. along the tet height sqrt(2/3) s
    P = np.array(REG) + 1.0
    s = np.linalg.norm(P[0] - P[1])
    tet = np.array([[0, 1, 2, 3]], np.int32)
    deg, ball, _ = O.vertex_ball(tet, 5, 32)                  # vertex 4 is isolated
    crd = np.vstack([P, [9.0, 9.0, 9.0]])
    new, item, res = O.face_pull64(tet, deg, ball, np.hstack([crd, np.zeros((5, 1))]), 1)
    for j in range(4):
        f = np.delete(P, j, axis=0).mean(axis=0)
        assert np.allclose(item[0, j, :3], f, atol=1e-14)
        assert np.allclose(new[j], P[j] + 0.25 * (f - P[j]), atol=1e-14)
        assert res[j] == pytest.approx(0.0625 * (2.0 / 3.0) * s * s, rel=1e-12)
    assert item[0, 0, 3] == pytest.approx(s ** 3 / (6 * math.sqrt(2)), rel=1e-12)   # regular tet volume
    assert np.array_equal(new[4], crd[4]) and res[4] == 0.0

    # two tets on face (a, b, c): apex d of a regular tet, apex e half as far
    # on the other side (half the volume).  The shared vertex a takes the
    # 2:1 volume-weighted mean of the centroids of faces (b, c, d) and (c, b, e)
    a, b, c, d = (np.array(x) for x in REG)
    ctr = (a + b + c) / 3.0
    e = ctr - 0.5 * (d - ctr)
    P = np.array([a, b, c, d, e])
    tet = np.array([[0, 1, 2, 3], [0, 2, 1, 4]], np.int32)
    deg, ball, _ = O.vertex_ball(tet, 5, 32)
    new, item, res = O.face_pull64(tet, deg, ball, np.hstack([P, np.zeros((5, 1))]), 1)
    assert item[1, 0, 3] == pytest.approx(0.5 * item[0, 0, 3], rel=1e-12)
    f1, f2 = (b + c + d) / 3.0, (c + b + e) / 3.0
    target = (2 * f1 + f2) / 3.0
    assert np.allclose(new[0], a + 0.25 * (target - a), atol=1e-14)
    # the apexes see one tet each: the centroid of (a, b, c)
    assert np.allclose(new[3], d + 0.25 * (ctr - d), atol=1e-14)
    assert np.allclose(new[4], e + 0.25 * (ctr - e), atol=1e-14)
    # an inverted tet weighs by its absolute volume
    inv = np.array([[0, 2, 1, 3]], np.int32)
    _, item2, _ = O.face_pull64(inv, *O.vertex_ball(inv, 5, 32)[:2], np.hstack([P, np.zeros((5, 1))]), 1)
    assert item2[0, 0, 3] == pytest.approx(item[0, 0, 3], rel=1e-12)


def test_bsr_spmv_f32_exact_small_integers():
    # small-integer values and x: every FP32 product and sum is exact, so the
    # FP32 product must equal the exact integer product, with non-symmetric
    # blocks (a transposed block or swapped row/column shows)
    rng = np.random.default_rng(4)
    for b in (1, 3, 4, 5):
        nb = 30
        A = sp.random(nb, nb, density=0.2, random_state=b, format="csr")
        A = (A + sp.eye(nb)).tocsr()
        rowptr, col = A.indptr.astype(np.int32), A.indices.astype(np.int32)
        val = rng.integers(-8, 9, size=(len(col), b, b)).astype(np.float32)
        x = rng.integers(-8, 9, size=nb * b).astype(np.float32)
        y = np.empty(nb * b, np.float32)
        O.lib().orc_bsr_spmv_f32(nb, b, O.p(rowptr), O.p(col), O.p(val), O.p(x), O.p(y))
        want = np.zeros(nb * b, np.int64)
        for r in range(nb):
            for kk in range(rowptr[r], rowptr[r + 1]):
                want[r * b:(r + 1) * b] += val[kk].astype(np.int64) @ x[col[kk] * b:(col[kk] + 1) * b].astype(np.int64)
        assert np.array_equal(y.astype(np.int64), want), b


def test_bsr_spmv_f32_within_rounding_of_f64():
    # FP32 accumulation of a row with m terms: |y32 - y64| <= gamma_m sum |a||x|
    rng = np.random.default_rng(5)
    nb, b = 200, 3
    A = (sp.random(nb, nb, density=0.08, random_state=2, format="csr") + sp.eye(nb)).tocsr()
    rowptr, col = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    val = rng.normal(size=(len(col), b, b)).astype(np.float32)
    x = rng.normal(size=nb * b).astype(np.float32)
    y = np.empty(nb * b, np.float32)
    O.lib().orc_bsr_spmv_f32(nb, b, O.p(rowptr), O.p(col), O.p(val), O.p(x), O.p(y))
    y64 = O.bsr_spmv64(rowptr, col, val, x.astype(np.float64), b)
    absv = O.bsr_spmv64(rowptr, col, np.abs(val), np.abs(x).astype(np.float64), b)
    m = (np.diff(rowptr) * b).repeat(b)
    u = np.finfo(np.float32).eps / 2
    assert np.all(np.abs(y - y64) <= (m * u / (1 - m * u)) * absv + 1e-30)


def test_bsr_spmv_vs_scipy():
    rng = np.random.default_rng(9)
    nb, b = 40, 3
    A = sp.random(nb, nb, density=0.15, random_state=1, format="csr")
    A = (A + sp.eye(nb)).tocsr()
    rowptr = A.indptr.astype(np.int32)
    col = A.indices.astype(np.int32)
    val = rng.normal(size=(len(col), b, b)).astype(np.float32)
    x = rng.normal(size=nb * b)
    y = O.bsr_spmv64(rowptr, col, val, x, b)
    M = sp.bsr_matrix((val.astype(np.float64), col, rowptr), shape=(nb * b, nb * b))
    assert np.allclose(y, M @ x, rtol=1e-12, atol=1e-12)
