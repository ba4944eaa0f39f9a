"""Python binding of libgml.so, the MI355X-native GMlib kernel engine.

The product is the C ABI declared in include/gmlib3.h (the reference GMlib
interface, sources/gmlib3.h:77-123) and include/gmlx.h (extensions).  This
module loads the in-tree libgml.so with ctypes and exposes every entry point
under its C name, plus a few numpy conveniences used by tests and bench.py.
It never falls back to anything else: if libgml.so is missing or fails to
load, importing the binding raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the library to load: libgml.so beside this file, or GMLIB_AMD_LIB (a checking build)
LIB_PATH = os.environ.get("GMLIB_AMD_LIB") or os.path.join(HERE, "libgml.so")

# enums and flags (sources/gmlib3.h:46-70)
GmlMaxDat, GmlMaxMat, GmlMaxVec, GmlMaxKrn = 100, 10, 10, 100
GmlRefFlag, GmlReadMode, GmlWriteMode, GmlVoyeurs, GmlManual = 1, 2, 4, 8, 16
(GmlVertices, GmlEdges, GmlTriangles, GmlQuadrilaterals,
 GmlTetrahedra, GmlPyramids, GmlPrisms, GmlHexahedra, GmlMaxEleTyp) = range(9)
(GmlInt, GmlInt2, GmlInt4, GmlInt8, GmlInt16,
 GmlFlt, GmlFlt2, GmlFlt4, GmlFlt8, GmlFlt16,
 GmlDbl, GmlDbl2, GmlDbl4, GmlDbl8, GmlDbl16,
 GmlByt, GmlByt2, GmlByt4, GmlByt8, GmlByt16, GmlMaxOclTyp) = range(21)
GmlMin, GmlMax, GmlSum, GmlL0, GmlL1, GmlL2, GmlLinf, GmlMaxRed = range(8)
(GmlxBuiltinQuality, GmlxBuiltinBallMin, GmlxBuiltinScatter, GmlxBuiltinGather, GmlxBuiltinSweep,
 GmlxBuiltinSpMV, GmlxBuiltinCgSpMV, GmlxBuiltinCgUpdate, GmlxBuiltinBatch) = range(9)
GmlxMaxBatch = 16
GmlxJobQuality, GmlxJobSmooth, GmlxJobCG = range(3)

_S = C.c_size_t
_I = C.c_int
_D = C.c_double
_F = C.c_float
_P = C.c_void_p
_CP = C.c_char_p
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# (restype, argtypes) of every non-variadic entry point
_SIGS = {
    "GmlInit": (_S, [_I]),
    "GmlStop": (None, [_S]),
    "GmlListGPU": (None, []),
    "GmlNewParameters": (_P, [_S, _I, _CP]),
    "GmlNewMeshData": (_I, [_S, _I, _I]),
    "GmlNewSolutionData": (_I, [_S, _I, _I, _I, _CP]),
    "GmlNewLinkData": (_I, [_S, _I, _I, _I, _CP]),
    "GmlNewMatrix": (_I, [_S, _I, _I, _I, _P, _P, _P, _I]),
    "GmlNewVector": (_I, [_S, _I, _I, _P, _I]),
    "GmlFreeData": (_I, [_S, _I]),
    "GmlSetCompilerOptions": (None, [_S, _CP]),
    "GmlLaunchKernel": (_I, [_S, _I]),
    "GmlReduceVector": (_I, [_S, _I, _I, _DP]),
    "GmlGetMemoryUsage": (_S, [_S]),
    "GmlGetMemoryTransfer": (_S, [_S]),
    "GmlGetMemoryAccess": (_F, [_S]),
    "GmlGetFlops": (_F, [_S]),
    "GmlDebugOn": (None, [_S]),
    "GmlDebugOff": (None, [_S]),
    "GmlExtractEdges": (_I, [_S]),
    "GmlExtractFaces": (_I, [_S]),
    "GmlSetNeighbours": (_I, [_S, _I]),
    "GmlCheckFP64": (_I, [_S]),
    "GmlGetMeshInfo": (_I, [_S, _I, _IP, _IP]),
    "GmlGetLinkInfo": (_I, [_S, _I, _I, _IP, _IP, _IP, _IP]),
    "GmlSetDataBlock": (_I, [_S, _I, _I, _I, _P, _P, _P, _P]),
    "GmlGetDataBlock": (_I, [_S, _I, _I, _I, _P, _P, _P, _P]),
    "GmlGetKernelRunTime": (_D, [_S, _I]),
    "GmlGetReduceRunTime": (_D, [_S, _I]),
    "GmlGetWallClock": (_D, []),
    "GmlUploadParameters": (_I, [_S]),
    "GmlDownloadParameters": (_I, [_S]),
    "GmlEvaluateNumbering": (_F, [_S]),
    "GmlIncludeUserToolkit": (None, [_S, _CP]),
    "GmlMultMatVec": (_I, [_S, _I, _I, _I]),
    "GmlMultDiagMatVec": (_I, [_S, _I, _I, _I]),
    "GmlAddVec3": (_I, [_S, _I, _I, _I, _I]),
    "GmlScaleVec": (_I, [_S, _I, _DP]),
    "GmlNormVec": (_I, [_S, _I, _I, _DP]),
    "GmlxInitHostOnly": (_S, []),
    "GmlxCompileKernelv": (_I, [_S, _CP, _CP, _I, _I, _IP]),
    "GmlxGetKernelSource": (_CP, [_S, _I]),
    "GmlxGetKernelPairSource": (_CP, [_S, _I]),
    "GmlxGetKernelRows": (_I, [_S, _I]),
    "GmlxGetKernelForm": (_I, [_S, _I]),
    "GmlxGetTilePlanInfo": (_I, [_S, _I, C.POINTER(C.c_longlong)]),
    "GmlxGetKernelTileSource": (_CP, [_S, _I]),
    "GmlxSetKernelShape": (_I, [_S, _I, _I, _I, _I]),
    "GmlxGetKernelShape": (_I, [_S, _I, _IP, _IP]),
    "GmlxGetDevicePointer": (_P, [_S, _I, C.POINTER(C.c_size_t)]),
    "GmlxSync": (_I, [_S]),
    "GmlxGetStream": (_P, [_S]),
    "GmlxInvalidateHost": (_I, [_S, _I]),
    "GmlxBuildLink": (_I, [_S, _I, _I]),
    "GmlxSetBuiltinItems": (_I, [_S, _I, _I]),
    "GmlxGetLinkData": (_I, [_S, _I, _I]),
    "GmlxGetUplinkTable": (_I, [_S, _I, _I, _P, _P, _I]),
    "GmlxCubeMesh": (_I, [_I, _D, C.c_ulonglong, _I, _P, _P, _P]),
    "GmlxSetVector": (_I, [_S, _I, _P]),
    "GmlxGetVector": (_I, [_S, _I, _P]),
    "GmlxFreeVector": (_I, [_S, _I]),
    "GmlxFreeMatrix": (_I, [_S, _I]),
    "GmlxGetMatrixInfo": (_I, [_S, _I, C.POINTER(C.c_longlong)]),
    "GmlxDotVec": (_I, [_S, _I, _I, _DP]),
    "GmlxReduceVectorNoAlloc": (_I, [_S, _I, _I, _DP]),
    "GmlxSolveCG": (_I, [_S, _I, _I, _I, _D, _I, _I, _DP, _IP]),
    "GmlxTetQuality": (_I, [_S, _I]),
    "GmlxBallMin": (_I, [_S, _I, _I]),
    "GmlxSmoothSweeps": (_I, [_S, _I, _I, _D]),
    "GmlxSetBuiltinShape": (_I, [_S, _I, _I, _I]),
    "GmlxGetBuiltinRunTime": (_D, [_S, _I, C.POINTER(C.c_longlong)]),
    "GmlxSetSmoothMode": (_I, [_S, _I]),
    "GmlxGetSweepPlanInfo": (_I, [_S, C.POINTER(C.c_longlong)]),
    "GmlxNewBatch": (_I, [_S, _I, _P, _P, _P, _P]),
    "GmlxBatchTune": (_I, [_S, _I, _IP, _IP]),
    "GmlxBatchSetShape": (_I, [_S, _I, _I, _I]),
    "GmlxBatchQuality": (_I, [_S, _I, _P, _P]),
    "GmlxGetBatchField": (_I, [_S, _I, _P, _P]),
    "GmlxGetBatchInfo": (_I, [_S, _I, C.POINTER(C.c_longlong)]),
    "GmlxGetBatchTuning": (_I, [_S, _I, _DP]),
    "GmlxFreeBatch": (_I, [_S, _I]),
    "GmlxEdgeLaplacian3": (C.c_longlong, [_I, _I, _P, C.c_ulonglong, _D, _P, _P, _P, C.c_longlong]),
    "GmlxServeStart": (_S, [_I, _D, _I, _I]),
    "GmlxServeSubmit": (C.c_longlong, [_S, _I, _I, C.c_ulonglong, _D]),
    "GmlxServeWait": (_I, [_S, C.c_longlong, _DP, _DP]),
    "GmlxServeAddMesh": (C.c_longlong, [_S, _I, _I, _P, _P, _P]),
    "GmlxServeSubmitMesh": (C.c_longlong, [_S, _I, C.c_longlong, _D]),
    "GmlxServeRemoveMesh": (_I, [_S, C.c_longlong]),
    "GmlxServeInfo": (_I, [_S, _DP]),
    "GmlxServeStop": (None, [_S]),
    "GmlxSetStaging": (_I, [_S, _I]),
    "GmlxHilbertRenumber": (_I, [_I, _P, _P, _I, _I, _P, _P, _P, _P, _I]),
}

# variadic entry points: no argtypes (callers pass ctypes values)
_VARIADIC = {"GmlSetDataLine": _I, "GmlGetDataLine": _I, "GmlCompileKernel": _I, "GmlImportMesh": _I,
             "GmlExportSolution": _I}

_lib = None


def lib():
    """The loaded libgml.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgml.so is not built (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
        for name, res in _VARIADIC.items():
            getattr(_lib, name).restype = res
    return _lib


def set_line(h, idx, lin, *args):
    """GmlSetDataLine through its C varargs (handle and indices typed here;
    the varargs must already be ctypes values: c_double, c_int, c_void_p)."""
    return lib().GmlSetDataLine(C.c_size_t(h), C.c_int(idx), C.c_int(lin), *args)


def get_line(h, idx, lin, *args):
    """GmlGetDataLine through its C varargs (pointers as ctypes byref/c_void_p)."""
    return lib().GmlGetDataLine(C.c_size_t(h), C.c_int(idx), C.c_int(lin), *args)


def declared_symbols():
    """Every entry point declared by include/gmlib3.h and include/gmlx.h."""
    return sorted(list(_SIGS) + list(_VARIADIC))


def __getattr__(name):
    if name.startswith("Gml"):
        return getattr(lib(), name)
    raise AttributeError(name)


def ptr(a):
    """Address of a contiguous numpy array as c_void_p."""
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# numpy conveniences
# ---------------------------------------------------------------------------

def cube_mesh(n, jitter=0.2, seed=1, hilbert=True):
    """(crd float32 [V,4], ref int32 [V], tet int32 [T,4]) of the n^3-cell
    Kuhn cube mesh built by GmlxCubeMesh."""
    nv = (n + 1) ** 3
    nt = 6 * n ** 3
    crd = np.empty((nv, 4), np.float32)
    ref = np.empty(nv, np.int32)
    tet = np.empty((nt, 4), np.int32)
    if not lib().GmlxCubeMesh(n, float(jitter), int(seed), 1 if hilbert else 0, ptr(crd), ptr(ref), ptr(tet)):
        raise RuntimeError("GmlxCubeMesh failed")
    return crd, ref, tet


def edge_laplacian3(nver, tet, seed, shift=0.01):
    """(rowptr, col, val[nnzb, 3, 3]) of GmlxEdgeLaplacian3 for a tet mesh."""
    tet = np.ascontiguousarray(tet, np.int32)
    rowptr = np.empty(nver + 1, np.int32)
    nnz = lib().GmlxEdgeLaplacian3(nver, len(tet), ptr(tet), seed, shift, ptr(rowptr), None, None, 0)
    if nnz < 0:
        raise ValueError("GmlxEdgeLaplacian3: invalid mesh")
    col = np.empty(nnz, np.int32)
    val = np.empty((nnz, 3, 3), np.float32)
    if lib().GmlxEdgeLaplacian3(nver, len(tet), ptr(tet), seed, shift, ptr(rowptr), ptr(col), ptr(val), nnz) != nnz:
        raise RuntimeError("GmlxEdgeLaplacian3 failed")
    return rowptr, col, val


def concat_meshes(meshes):
    """(nver int32[M], ntet int32[M], crd float32[sum V, 4], tet int32[sum T, 4])
    of a list of (crd, ref, tet) meshes, tets keeping their local indices:
    the input layout of GmlxNewBatch."""
    nver = np.array([len(m[0]) for m in meshes], np.int32)
    ntet = np.array([len(m[2]) for m in meshes], np.int32)
    crd = np.ascontiguousarray(np.concatenate([m[0] for m in meshes]), np.float32)
    tet = np.ascontiguousarray(np.concatenate([m[2] for m in meshes]), np.int32)
    return nver, ntet, crd, tet


class Gml:
    """One GmlInit handle with numpy helpers (thin: every call goes to libgml)."""

    def __init__(self, dev=0, host_only=False):
        self.L = lib()
        self.h = self.L.GmlxInitHostOnly() if host_only else self.L.GmlInit(dev)
        if not self.h:
            raise RuntimeError("GmlInit(%d) failed" % dev)
        self._keep = []

    def close(self):
        if self.h:
            self.L.GmlStop(self.h)
            self.h = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def keep(self, obj):
        """Keep a buffer alive for the handle's lifetime (the library keeps
        caller strings by pointer, as the reference does)."""
        self._keep.append(obj)
        return obj

    def cstr(self, s):
        b = C.create_string_buffer(s.encode() if isinstance(s, str) else s)
        return self.keep(b)

    # --- data -----------------------------------------------------------
    def new_parameters(self, size, src):
        p = self.L.GmlNewParameters(self.h, size, self.cstr(src))
        if not p:
            raise RuntimeError("GmlNewParameters failed")
        return p

    def set_mesh(self, crd, ref, tet, vref=None):
        """Create vertices and tets and upload them with GmlSetDataBlock."""
        nv, nt = len(crd), len(tet)
        vi = self.L.GmlNewMeshData(self.h, GmlVertices, nv)
        ti = self.L.GmlNewMeshData(self.h, GmlTetrahedra, nt)
        if not vi or not ti:
            raise RuntimeError("GmlNewMeshData failed")
        crd = np.ascontiguousarray(crd, np.float32)
        ref = np.ascontiguousarray(ref, np.int32)
        tet = np.ascontiguousarray(tet, np.int32)
        tref = np.zeros(nt, np.int32)
        w = crd.shape[1]
        ok = self.L.GmlSetDataBlock(self.h, GmlVertices, 0, nv - 1, ptr(crd),
                                    C.c_void_p(crd.ctypes.data + (nv - 1) * w * 4),
                                    ptr(ref), C.c_void_p(ref.ctypes.data + (nv - 1) * 4))
        ok &= self.L.GmlSetDataBlock(self.h, GmlTetrahedra, 0, nt - 1, ptr(tet),
                                     C.c_void_p(tet.ctypes.data + (nt - 1) * 16),
                                     ptr(tref), C.c_void_p(tref.ctypes.data + (nt - 1) * 4))
        if not ok:
            raise RuntimeError("GmlSetDataBlock failed")
        return vi, ti

    def solution(self, typ, n, itm, name):
        i = self.L.GmlNewSolutionData(self.h, typ, n, itm, self.cstr(name))
        if not i:
            raise RuntimeError("GmlNewSolutionData failed")
        return i

    def compile(self, src, name, typ, args):
        """args: list of (DatIdx, flags, LnkIdx)."""
        tri = (C.c_int * (3 * len(args)))(*[int(x) for a in args for x in a])
        k = self.L.GmlxCompileKernelv(self.h, self.cstr(src), self.cstr(name), typ, len(args), tri)
        if not k:
            raise RuntimeError("GmlCompileKernel(%s) failed" % name)
        return k

    def launch(self, k):
        r = self.L.GmlLaunchKernel(self.h, k)
        if r != 1:
            raise RuntimeError("GmlLaunchKernel(%d) returned %d" % (k, r))

    def reduce(self, idx, op):
        v = C.c_double(0.)
        r = self.L.GmlReduceVector(self.h, idx, op, C.byref(v))
        if r != 1:
            raise RuntimeError("GmlReduceVector returned %d" % r)
        return v.value

    def download(self, idx, dtype, shape):
        """Whole data item copied to a numpy array (GmlGetDataLine line 0
        downloads the item; the other lines read the mirror)."""
        n = shape[0]
        out = np.empty(shape, dtype)
        row = out.reshape(n, -1)
        for i in range(n):
            if not get_line(self.h, idx, i, C.c_void_p(row[i].ctypes.data)):
                raise RuntimeError("GmlGetDataLine failed")
        return out

    def download_raw(self, idx, dtype, shape):
        """Device contents of a data item via its device pointer (fast path)."""
        import ctypes
        nb = C.c_size_t(0)
        dp = self.L.GmlxGetDevicePointer(self.h, idx, C.byref(nb))
        out = np.empty(shape, dtype)
        assert out.nbytes <= nb.value
        self.L.GmlxSync(self.h)
        _hip_d2h(out, dp)
        return out

    def upload(self, idx, a):
        """Copy a numpy array into a data item's device memory (its byte
        size at most the item's) and mark the host mirror stale."""
        a = np.ascontiguousarray(a)
        nb = C.c_size_t(0)
        dp = self.L.GmlxGetDevicePointer(self.h, idx, C.byref(nb))
        if not dp or a.nbytes > nb.value:
            raise RuntimeError("upload of %d bytes into data %d (%d bytes)" % (a.nbytes, idx, nb.value))
        self.L.GmlxSync(self.h)
        _hip_h2d(dp, a)
        self.L.GmlxInvalidateHost(self.h, idx)

    def vertices(self, nv):
        crd = np.zeros((nv, 4), np.float32)   # GmlGetDataBlock writes x, y, z; w stays 0
        ref = np.empty(nv, np.int32)
        ok = self.L.GmlGetDataBlock(self.h, GmlVertices, 0, nv - 1, ptr(crd),
                                    C.c_void_p(crd.ctypes.data + (nv - 1) * 16),
                                    ptr(ref), C.c_void_p(ref.ctypes.data + (nv - 1) * 4))
        if not ok:
            raise RuntimeError("GmlGetDataBlock failed")
        return crd, ref

    def uplink_table(self, src, dst, nsrc, width):
        deg = np.empty(nsrc, np.int32)
        ball = np.empty((nsrc, width), np.int32)
        if not self.L.GmlxGetUplinkTable(self.h, src, dst, ptr(deg), ptr(ball), width):
            raise RuntimeError("GmlxGetUplinkTable failed")
        return deg, ball

    def link_info(self, src, dst):
        v = [C.c_int(0) for _ in range(4)]
        if not self.L.GmlGetLinkInfo(self.h, src, dst, *[C.byref(x) for x in v]):
            return None
        return tuple(x.value for x in v)

    def kernel_time(self, k):
        return self.L.GmlGetKernelRunTime(self.h, k)


_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    return _hip


def _hip_d2h(out, dptr):
    r = _hiprt().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(dptr), out.nbytes, 2)
    if r != 0:
        raise RuntimeError("hipMemcpy failed: %d" % r)


def _hip_h2d(dptr, a):
    a = np.ascontiguousarray(a)
    r = _hiprt().hipMemcpy(C.c_void_p(dptr), C.c_void_p(a.ctypes.data), a.nbytes, 1)
    if r != 0:
        raise RuntimeError("hipMemcpy failed: %d" % r)
