"""ctypes view of oracle/liboracle.so, the CPU restatement used as the checker.

Test infrastructure only: tests, __graft_entry__.smoke() and bench.py's CPU
baseline load it; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_F = C.c_float


class F4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = C.CDLL(LIB)
        sigs = {
            "orc_vertex_ball": (_I, [_I, _I, _I, _I, _P, _I, _P, _P]),
            "orc_get_cnt_vec": (None, [_I, _P, _P]),
            "orc_evaluate_numbering": (_F, [_I, _I, _P]),
            "orc_meanratio_f64": (_D, [_P, _P, _P, _P]),
            "orc_meanratio_f32": (_F, [F4, F4, F4, F4]),
            "orc_ref_compute_quality_f64": (_D, [_P, _P, _P, _P]),
            "orc_caltetqal_f64": (_D, [_P, _P, _P, _P]),
            "orc_caltriqal_f64": (_D, [_P, _P, _P]),
            "orc_tet_meanratio_f64_all": (None, [_I, _P, _P, _P]),
            "orc_tet_meanratio_f32_all": (None, [_I, _P, _P, _P]),
            "orc_ball_min": (None, [_I, _P, _P, _I, _P, _P]),
            "orc_reduce": (_D, [_I, C.c_longlong, _P]),
            "orc_smooth_sweep_f32": (None, [_I, _I, _P, _P, _P, _P, _I, _P, _P, _F]),
            "orc_smooth_sweep_f64": (None, [_I, _I, _P, _P, _P, _P, _I, _P, _P, _D]),
            "orc_ref_nodesmooth_sweep_f64": (None, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
            "orc_face_pull_sweep_f64": (None, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
            "orc_bsr_spmv_f64": (None, [_I, _I, _P, _P, _P, _P, _P]),
            "orc_bsr_spmv_f32": (None, [_I, _I, _P, _P, _P, _P, _P]),
            "orc_extract_items": (C.c_longlong, [_I, _I, _P, _P, _P, _P, _P, _I, _P, C.c_longlong, _P, _P,
                                                  C.c_longlong]),
            "orc_bw_direct_tets": (None, [_I, _P, _P]),
            "orc_bw_indirect_tets": (None, [_I, _P, _P, _P]),
            "orc_bw_vertices_ball": (None, [_I, _P, _P, _I, _P, _P]),
            "orc_num_threads": (_I, []),
        }
        for n, (r, a) in sigs.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def p(a):
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return C.c_void_p(a.ctypes.data)


def d3(x):
    return np.ascontiguousarray(x, np.float64)


def meanratio64(a, b, c, d):
    a, b, c, d = (d3(x) for x in (a, b, c, d))
    return lib().orc_meanratio_f64(p(a), p(b), p(c), p(d))


def vertex_ball(tet, nver, width):
    tet = np.ascontiguousarray(tet, np.int32)
    deg = np.zeros(nver, np.int32)
    ball = np.zeros((nver, width), np.int32)
    md = lib().orc_vertex_ball(nver, len(tet), tet.shape[1], tet.shape[1], p(tet), width, p(deg), p(ball))
    return deg, ball, md


def tet_quality64(tet, crd):
    q = np.empty(len(tet), np.float64)
    lib().orc_tet_meanratio_f64_all(len(tet), p(np.ascontiguousarray(tet, np.int32)),
                                    p(np.ascontiguousarray(crd, np.float32)), p(q))
    return q


def tet_quality32(tet, crd):
    q = np.empty(len(tet), np.float32)
    lib().orc_tet_meanratio_f32_all(len(tet), p(np.ascontiguousarray(tet, np.int32)),
                                    p(np.ascontiguousarray(crd, np.float32)), p(q))
    return q


def ball_min(deg, ball, val):
    out = np.empty(len(deg), np.float32)
    lib().orc_ball_min(len(deg), p(deg), p(ball), ball.shape[1], p(np.ascontiguousarray(val, np.float32)), p(out))
    return out


def reduce(op, x):
    x = np.ascontiguousarray(x, np.float32)
    return lib().orc_reduce(op, len(x), p(x))


def smooth32(tet, ref, deg, ball, crd, nsweep, omega):
    crd = np.array(crd, np.float32, copy=True)
    bary = np.empty((len(tet), 4), np.float32)
    for _ in range(nsweep):
        lib().orc_smooth_sweep_f32(len(crd), len(tet), p(tet), p(ref), p(deg), p(ball), ball.shape[1],
                                   p(crd), p(bary), omega)
    return crd


def smooth64(tet, ref, deg, ball, crd, nsweep, omega):
    c = np.ascontiguousarray(np.asarray(crd, np.float64)[:, :3])
    bary = np.empty((len(tet), 3), np.float64)
    for _ in range(nsweep):
        lib().orc_smooth_sweep_f64(len(c), len(tet), p(tet), p(ref), p(deg), p(ball), ball.shape[1],
                                   p(c), p(bary), omega)
    return c


def ref_nodesmooth64(tet, deg, ball, crd, nsweep):
    c = np.ascontiguousarray(np.asarray(crd, np.float64)[:, :3])
    opt = np.empty((len(tet), 4, 3), np.float64)
    res = np.empty(len(c), np.float64)
    for _ in range(nsweep):
        lib().orc_ref_nodesmooth_sweep_f64(len(c), len(tet), p(tet), p(deg), p(ball), ball.shape[1],
                                           p(c), p(opt), p(res))
    return c, opt, res


def face_pull64(tet, deg, ball, crd, nsweep):
    """nsweep sweeps of the volume-weighted opposite-face pull (FP64):
    (coordinates, items of the last sweep, residuals of the last sweep)."""
    c = np.ascontiguousarray(np.asarray(crd, np.float64)[:, :3])
    item = np.empty((len(tet), 4, 4), np.float64)
    res = np.zeros(len(c), np.float64)
    for _ in range(nsweep):
        lib().orc_face_pull_sweep_f64(len(c), len(tet), p(tet), p(deg), p(ball), ball.shape[1], p(c), p(item), p(res))
    return c, item, res


def bsr_spmv64(rowptr, col, val, x, b):
    y = np.empty(len(rowptr) - 1, dtype=np.float64).repeat(b)
    lib().orc_bsr_spmv_f64(len(rowptr) - 1, b, p(rowptr), p(col), p(val), p(np.ascontiguousarray(x, np.float64)), p(y))
    return y


# Local items of each element kind (reference tables ItmEdgNod and
# ItmFacNod, gmlib3.c:330-358), keyed by GmlTriangles..GmlHexahedra = 2..7.
EDGE_LOC = {
    2: [(0, 1), (1, 2), (2, 0)],
    3: [(0, 1), (1, 2), (2, 3), (3, 0)],
    4: [(0, 1), (1, 2), (2, 0), (3, 0), (3, 1), (3, 2)],
    5: [(0, 1), (3, 2), (0, 3), (1, 2), (0, 4), (1, 4), (2, 4), (3, 4)],
    6: [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (0, 3), (1, 4), (2, 5)],
    7: [(0, 1), (3, 2), (7, 6), (4, 5), (0, 3), (4, 7), (5, 6), (1, 2), (0, 4), (1, 5), (2, 6), (3, 7)],
}
FACE_LOC = {
    2: [(0, 1, 2)],
    3: [(0, 1, 2, 3)],
    4: [(1, 2, 3), (2, 0, 3), (3, 0, 1), (0, 2, 1)],
    5: [(0, 1, 4), (1, 2, 4), (2, 3, 4), (3, 0, 4), (3, 2, 1, 0)],
    6: [(0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (3, 5, 2, 0)],
    7: [(0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2), (0, 3, 2, 1), (4, 5, 6, 7)],
}


def extract_items(deg, tables, old=None, oldref=None, tabsiz=0):
    """Distinct items of degree deg (2 edges, 3 triangles, 4 quads) of the
    element tables [(kind, int32 [n, nnod])...], visited in order after the
    existing table old ([n, deg], refs oldref): (nodes [m, deg], refs [m])."""
    srcs = []
    if old is not None:
        srcs.append((np.ascontiguousarray(old, np.int32), [tuple(range(deg))]))
    for kind, t in tables:
        loc = EDGE_LOC[kind] if deg == 2 else [f for f in FACE_LOC[kind] if len(f) == deg]
        if loc and len(t):
            srcs.append((np.ascontiguousarray(t, np.int32), loc))
    if not srcs:
        return np.zeros((0, deg), np.int32), np.zeros(0, np.int32)
    ns = len(srcs)
    ne = np.array([len(t) for t, _ in srcs], np.int64)
    stride = np.array([t.shape[1] for t, _ in srcs], np.int32)
    nloc = np.array([len(l) for _, l in srcs], np.int32)
    loc = np.zeros((ns, 12, 4), np.int32)
    for s, (_, l) in enumerate(srcs):
        for j, f in enumerate(l):
            loc[s, j, :len(f)] = f
    ptrs = (C.c_void_p * ns)(*[t.ctypes.data for t, _ in srcs])
    cap = int((ne * nloc).sum())
    nodes = np.zeros((cap, deg), np.int32)
    refs = np.zeros(cap, np.int32)
    oref = np.ascontiguousarray(oldref if oldref is not None else np.zeros(len(old) if old is not None else 0),
                                np.int32)
    m = lib().orc_extract_items(deg, ns, p(ne), p(stride), p(nloc), p(loc), C.cast(ptrs, C.c_void_p),
                                1 if old is not None else 0, p(oref), int(tabsiz), p(nodes), p(refs), cap)
    if m < 0:
        raise MemoryError("orc_extract_items")
    return nodes[:m].copy(), refs[:m].copy()


# ---------------------------------------------------------------------------
# The reference's hash links, restated in pure Python for small meshes:
# GetItmNod (gmlib3.c:1567-1620), CalHshKey (1627-1638), AddHsh (1645-1674),
# GetHsh (1681-1721), used by NewBallData (743-1168) for shells and
# downlinks and by GmlSetNeighbours (4178-4267).  The row orders come out of
# the chain walk itself (head insertion, walk from the head).
# ---------------------------------------------------------------------------

# NmbTpoLnk[element kind][item kind] (gmlib3.c:316-325) for the kinds used here
TPO = {(1, 0): 2, (2, 1): 3, (3, 1): 4, (4, 0): 4, (4, 1): 6, (4, 2): 4, (5, 1): 8, (7, 0): 8, (7, 1): 12,
       (7, 3): 6, (1, 1): 2, (2, 2): 3, (3, 3): 4, (4, 4): 4, (7, 7): 6}
NGB_TYP = {1: 0, 2: 1, 3: 1, 4: 2, 7: 3}           # NgbTyp (gmlib3.c:327)
ITEM_NODES = {0: 1, 1: 2, 2: 3, 3: 4}


def item_nodes(ele, ele_kind, itm_kind, j):
    """GetItmNod: the sorted nodes of local item j of an element."""
    if itm_kind == 0:
        return (int(ele[j]),)
    if itm_kind == 1:
        loc = EDGE_LOC[ele_kind][j] if ele_kind >= 2 else (0, 1)
        return tuple(sorted(int(ele[k]) for k in loc))
    f = FACE_LOC[ele_kind][j]
    return tuple(sorted(int(ele[k]) for k in f[:ITEM_NODES[itm_kind]]))


class RefHash:
    W = (3, 5, 7, 11)

    def __init__(self, tabsiz, datlen):
        self.tabsiz, self.datlen = max(1, tabsiz), datlen
        self.head = {}
        self.buc = [None]

    def key(self, nodes):
        if self.datlen == 1:
            return nodes[0]
        return sum(w * n for w, n in zip(self.W, nodes)) % self.tabsiz

    def add(self, nodes, e, j):
        k = self.key(nodes)
        self.buc.append((tuple(nodes), e, j, self.head.get(k, 0)))
        self.head[k] = len(self.buc) - 1

    def get(self, nodes):
        """codes (e << 4) | j of the matches, in chain order"""
        out, b = [], self.head.get(self.key(nodes), 0)
        while b:
            nod, e, j, nxt = self.buc[b]
            if nod == tuple(nodes):
                out.append((e << 4) | j)
            b = nxt
        return out


def ref_shells(src, src_kind, dst, dst_kind):
    """Uplink rows (edge or face -> elements) in the chain order."""
    h = RefHash(len(src), ITEM_NODES[src_kind])
    for i, e in enumerate(dst):
        for j in range(TPO[(dst_kind, src_kind)]):
            h.add(item_nodes(e, dst_kind, src_kind, j), i, j)
    return [h.get(item_nodes(s, src_kind, src_kind, 0)) for s in src]


def ref_downlink(src, src_kind, dst, dst_kind, width):
    """Downlink table: the first chain match's entity index, 0 if none."""
    h = RefHash(len(dst), ITEM_NODES[dst_kind])
    for i, d in enumerate(dst):
        h.add(item_nodes(d, dst_kind, dst_kind, 0), i, 0)
    out = np.zeros((len(src), width), np.int32)
    for i, e in enumerate(src):
        for j in range(TPO[(src_kind, dst_kind)]):
            m = h.get(item_nodes(e, src_kind, dst_kind, j))
            if m:
                out[i, j] = m[0] >> 4
    return out


def ref_neighbours(tab, kind, width):
    """GmlSetNeighbours' TetNgb rows: per item, the last chain match of
    another element, 0 if none."""
    ng = NGB_TYP[kind]
    ni = TPO[(kind, ng)]
    h = RefHash(len(tab) * TPO[(kind, kind)] // 2, ITEM_NODES[ng])
    for i, e in enumerate(tab):
        for j in range(ni):
            h.add(item_nodes(e, kind, ng, j), i, j)
    out = np.zeros((len(tab), width), np.int32)
    for i, e in enumerate(tab):
        for j in range(ni):
            for c in h.get(item_nodes(e, kind, ng, j)):
                if c >> 4 != i:
                    out[i, j] = c >> 4
    return out


# ---------------------------------------------------------------------------
# Cell-centered heat equation (the reference's examples/HeatCellCentered
# scheme: sol_ext.cl, grd_tet.cl, grd_ext.cl, flx_bal.cl, tim_int.cl, dt.cl),
# restated in numpy FP64.  Links from the mesh: the two tets of each
# triangle in descending index (the shell order; the second is absent on the
# boundary and reads as 0) and the triangle of each tet face, faces in the
# order opposite vertex 0, 1, 2, 3.
# ---------------------------------------------------------------------------

HEAT_FACES = ((1, 2, 3), (2, 0, 3), (3, 0, 1), (0, 2, 1))
CAL_TET_VOL = 0.166666          # the toolkit's CalTetVol factor (toolkit.cl)


def heat_links(nver, tet, tri):
    """(shell [F, 2] with -1 for none, down [T, 4])"""
    tet = np.asarray(tet, np.int64)
    tri = np.asarray(tri, np.int64)
    V = nver

    def key(a):
        a = np.sort(a, axis=1)
        return (a[:, 0] * V + a[:, 1]) * V + a[:, 2]

    ktri = key(tri)
    order = np.argsort(ktri, kind="stable")
    T = len(tet)
    down = np.zeros((T, 4), np.int64)
    shell = -np.ones((len(tri), 2), np.int64)
    for j, f in enumerate(HEAT_FACES):
        k = key(tet[:, list(f)])
        pos = np.searchsorted(ktri[order], k)
        assert (ktri[order][pos] == k).all()
        down[:, j] = order[pos]
    # tets of each triangle, descending
    t = np.repeat(np.arange(T), 4)
    f = down.ravel()
    o = np.lexsort((-t, f))
    f, t = f[o], t[o]
    first = np.r_[True, f[1:] != f[:-1]]
    shell[f[first], 0] = t[first]
    second = ~first
    shell[f[second], 1] = t[second]
    return shell, down


def heat_geometry(crd, tet):
    """area vectors (x2) of the four faces of each tet [T, 4, 3] and the volume as CalTetVol [T]"""
    c = np.asarray(crd, np.float64)[:, :3]
    tet = np.asarray(tet, np.int64)
    P = [c[tet[:, i]] for i in range(4)]
    nrm = np.stack([np.cross(P[b] - P[a], P[cc] - P[a]) for a, b, cc in HEAT_FACES], 1)
    vol = CAL_TET_VOL * np.einsum("ij,ij->i", np.cross(P[1] - P[0], P[2] - P[0]), P[3] - P[0])
    return nrm, vol


def heat_iterations(crd, tet, tri, triref, bc, niter, sol=None, dt=None):
    """niter explicit steps from sol (zeros by default): (sol, [residual per step], dt)"""
    shell, down = heat_links(len(crd), tet, tri)
    nrm, vol = heat_geometry(crd, tet)
    if dt is None:
        dt = float((0.3 * vol ** (2. / 3.)).min())
    u = np.zeros(len(tet)) if sol is None else np.asarray(sol, np.float64).copy()
    has2 = shell[:, 1] >= 0
    s1 = np.where(has2, shell[:, 1], 0)
    res = []
    for _ in range(niter):
        uf = 0.5 * (u[shell[:, 0]] + np.where(has2, u[s1], 0.))
        for kind, ref, val in bc:
            m = triref == ref
            uf[m] = val if kind == 0 else u[shell[m, 0]]
        g = 0.5 * np.einsum("tj,tjk->tk", uf[down], nrm) / vol[:, None]
        gf = 0.5 * (g[shell[:, 0]] + np.where(has2[:, None], g[s1], 0.))
        for kind, ref, val in bc:
            m = triref == ref
            gf[m] = g[shell[m, 0]] if kind == 0 else val
        rhs = 0.5 * np.einsum("tjk,tjk->t", gf[down], nrm) / vol
        u = u + dt * rhs
        res.append(float((rhs * rhs).sum()))
    return u, res, dt


def bw_direct_tets(tet):
    tet = np.ascontiguousarray(tet, np.int32)
    out = np.empty((len(tet), 4), np.float32)
    lib().orc_bw_direct_tets(len(tet), p(tet), p(out))
    return out


def bw_indirect_tets(tet, crd):
    tet = np.ascontiguousarray(tet, np.int32)
    crd = np.ascontiguousarray(crd, np.float32)
    out = np.empty((len(tet), 4), np.float32)
    lib().orc_bw_indirect_tets(len(tet), p(tet), p(crd), p(out))
    return out


def bw_vertices_ball(deg, ball, tdat):
    tdat = np.ascontiguousarray(tdat, np.float32)
    out = np.empty((len(deg), 4), np.float32)
    lib().orc_bw_vertices_ball(len(deg), p(deg), p(ball), ball.shape[1], p(tdat), p(out))
    return out
