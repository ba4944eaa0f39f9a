"""Cell-centered heat equation on a tet mesh through the GMlib C API: the
application the reference ships as examples/HeatCellCentered (HeatCellCentered.c),
driven the same way -- boundary patches in the parameter block, the volume
faces extracted after the boundary triangles (GmlExtractFaces), face values
and gradients on triangles through the triangle -> tet shells, cell
gradients and flux balances on tets through the tet -> triangle downlink,
explicit steps until the GmlReduceVector L2 residual has dropped by the
requested factor.  The kernel bodies are gmlib_amd/kernels/heat_*.cl.

heat_cube(n, ...) gives the input: a Kuhn cube mesh and its boundary
triangles with one ref per side (1: x = 0, 2: x = 1, 3: y = 0, 4: z = 0,
5: z = 1, 6: y = 1).
"""
import ctypes as C
import os

import numpy as np

from . import Gml, cube_mesh, ptr
from . import (GmlFlt, GmlFlt4, GmlL2, GmlMin, GmlReadMode, GmlRefFlag, GmlTetrahedra, GmlTriangles,
               GmlVertices, GmlWriteMode)

KDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernels")
FIXED, FLUX = 0, 1
# the reference example's setting: sides 4 and 5 held at 300 and 250, the
# other sides insulated (HeatCellCentered.c:91)
DEFAULT_BC = ((FLUX, 1, 0.), (FLUX, 2, 0.), (FLUX, 3, 0.), (FIXED, 4, 300.), (FIXED, 5, 250.), (FLUX, 6, 0.))
TET_FACES = ((1, 2, 3), (2, 0, 3), (3, 0, 1), (0, 2, 1))


def body(name):
    return open(os.path.join(KDIR, name + ".cl")).read()


class HeatParams(C.Structure):
    _fields_ = [("kind", C.c_int * 6), ("ref", C.c_int * 6), ("val", C.c_float * 6), ("dt", C.c_float)]


def heat_cube(n, jitter=0.2, seed=1):
    """(crd, vref, tet, tri, triref): a Kuhn cube mesh and its 12 n^2
    boundary triangles, ref = side number."""
    crd, vref, tet = cube_mesh(n, jitter, seed)
    f = np.concatenate([tet[:, list(t)] for t in TET_FACES])
    key = np.sort(f, axis=1).astype(np.int64)
    V = len(crd)
    k = (key[:, 0] * V + key[:, 1]) * V + key[:, 2]
    uk, first, cnt = np.unique(k, return_index=True, return_counts=True)
    tri = f[first[cnt == 1]]
    tri = tri[np.argsort(first[cnt == 1])]
    c = crd[tri, :3]
    side = np.zeros(len(tri), np.int32)
    for r, (axis, val) in {1: (0, 0.), 2: (0, 1.), 3: (1, 0.), 4: (2, 0.), 5: (2, 1.), 6: (1, 1.)}.items():
        side[(c[:, :, axis] == np.float32(val)).all(axis=1)] = r
    assert (side > 0).all()
    return crd, vref, tet, np.ascontiguousarray(tri, np.int32), side


class HeatSolver:
    """One GmlInit handle running the heat solver (the reference example's
    sequence of calls)."""

    def __init__(self, crd, vref, tet, tri, triref, bc=DEFAULT_BC, dev=0):
        self.g = g = Gml(dev)
        L, h = g.L, g.h
        par = L.GmlNewParameters(h, C.sizeof(HeatParams), g.cstr(body("heat_params")))
        if not par:
            raise RuntimeError("GmlNewParameters failed")
        self.par = C.cast(par, C.POINTER(HeatParams)).contents
        for p, (kind, ref, val) in enumerate(bc):
            self.par.kind[p], self.par.ref[p], self.par.val[p] = kind, ref, val
        self.ver, _ = g.set_mesh(crd, vref, tet)
        nt = len(tri)
        if not L.GmlNewMeshData(h, GmlTriangles, nt):
            raise RuntimeError("GmlNewMeshData(triangles) failed")
        tri = np.ascontiguousarray(tri, np.int32)
        triref = np.ascontiguousarray(triref, np.int32)
        if not L.GmlSetDataBlock(h, GmlTriangles, 0, nt - 1, ptr(tri), C.c_void_p(tri.ctypes.data + (nt - 1) * 12),
                                 ptr(triref), C.c_void_p(triref.ctypes.data + (nt - 1) * 4)):
            raise RuntimeError("GmlSetDataBlock(triangles) failed")
        # every face of the volume, the boundary triangles first with their refs
        self.nface = L.GmlExtractFaces(h)
        self.ntet = len(tet)
        n = C.c_int(0)
        idx = C.c_int(0)
        L.GmlGetMeshInfo(h, GmlTriangles, C.byref(n), C.byref(idx))
        self.tri = idx.value
        self.sol = g.solution(GmlTetrahedra, 1, GmlFlt, "SolTet")
        self.grd = g.solution(GmlTetrahedra, 1, GmlFlt4, "GrdTet")
        self.ext = g.solution(GmlTriangles, 1, GmlFlt, "SolExt")
        self.gext = g.solution(GmlTriangles, 1, GmlFlt4, "GrdExt")
        self.rhs = g.solution(GmlTetrahedra, 1, GmlFlt, "Rhs")
        self.dt = g.solution(GmlTetrahedra, 1, GmlFlt, "dt")
        R, W = GmlReadMode, GmlWriteMode
        self.k_face = g.compile(body("heat_face_value"), "heat_face_value", GmlTriangles,
                                [(self.tri, R | GmlRefFlag, 0), (self.sol, R, 0), (self.ext, W, 0)])
        self.k_grad = g.compile(body("heat_cell_gradient"), "heat_cell_gradient", GmlTetrahedra,
                                [(self.ver, R, 0), (self.ext, R, 0), (self.grd, W, 0)])
        self.k_fgrad = g.compile(body("heat_face_gradient"), "heat_face_gradient", GmlTriangles,
                                 [(self.tri, R | GmlRefFlag, 0), (self.grd, R, 0), (self.gext, W, 0)])
        self.k_bal = g.compile(body("heat_balance"), "heat_balance", GmlTetrahedra,
                               [(self.ver, R, 0), (self.gext, R, 0), (self.rhs, W, 0)])
        self.k_step = g.compile(body("heat_step"), "heat_step", GmlTetrahedra,
                                [(self.rhs, R, 0), (self.sol, R | W, 0)])
        self.k_dt = g.compile(body("heat_dt"), "heat_dt", GmlTetrahedra, [(self.ver, R, 0), (self.dt, W, 0)])
        # time step: the smallest cell's
        g.launch(self.k_dt)
        self.par.dt = g.reduce(self.dt, GmlMin)
        if L.GmlUploadParameters(h) != 1:
            raise RuntimeError("GmlUploadParameters failed")
        self.res0 = None

    def iterate(self, k):
        """k explicit steps; the L2 residual (sum of Rhs^2) of each."""
        g = self.g
        out = []
        for _ in range(k):
            for kr in (self.k_face, self.k_grad, self.k_fgrad, self.k_bal, self.k_step):
                g.launch(kr)
            out.append(g.reduce(self.rhs, GmlL2))
        if self.res0 is None and out:
            self.res0 = out[0]
        return out

    def solve(self, drop=1e-6, max_iter=100000, chunk=100):
        """Steps until the residual has dropped by drop (relative to the
        first step's), as the reference's loop; returns the step count."""
        n = 0
        while n < max_iter:
            r = self.iterate(min(chunk, max_iter - n))
            n += len(r)
            if r[-1] / self.res0 <= drop:
                break
        return n

    def temperature(self):
        return self.g.download_raw(self.sol, np.float32, (self.ntet,))

    def close(self):
        self.g.close()
