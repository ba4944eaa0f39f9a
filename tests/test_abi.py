"""The C-ABI boundary: libgml.so loads, exports every declared entry point,
and the host-side logic (data registry, line/block transfers, error codes)
behaves as the reference interface specifies.  Host-only handles: no GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import gmlib_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("gmlib3.h", "gmlx.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(Gmlx?[A-Z]\w*)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = header_functions()
    # the reference's 42 entry points plus its two libMeshb ones (GmlImportMesh, GmlExportSolution)
    assert len([n for n in names if not n.startswith("Gmlx")]) == 44
    lib = G.lib()
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(G.declared_symbols()) == names


def test_enum_values_match_reference_header():
    assert (G.GmlVertices, G.GmlTetrahedra, G.GmlHexahedra) == (0, 4, 7)
    assert (G.GmlInt, G.GmlFlt, G.GmlFlt4, G.GmlDbl, G.GmlByt16) == (0, 5, 7, 10, 19)
    assert (G.GmlMin, G.GmlSum, G.GmlL2, G.GmlLinf) == (0, 2, 5, 6)
    assert (G.GmlRefFlag, G.GmlReadMode, G.GmlWriteMode, G.GmlVoyeurs, G.GmlManual) == (1, 2, 4, 8, 16)


def test_line_and_block_roundtrip_host_only():
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi = L.GmlNewMeshData(h, G.GmlVertices, 3)
        ti = L.GmlNewMeshData(h, G.GmlTetrahedra, 2)
        assert vi and ti
        for i in range(3):
            assert G.set_line(h, vi, i, C.c_double(i + .5), C.c_double(2. * i), C.c_double(-i), C.c_int(7 * i))
        assert G.set_line(h, ti, 0, *[C.c_int(x) for x in (0, 1, 2, 0, 11)])
        assert G.set_line(h, ti, 1, *[C.c_int(x) for x in (2, 1, 0, 1, 12)])
        x, y, z = C.c_double(), C.c_double(), C.c_double()
        for i in (0, 2, 1):
            assert G.get_line(h, vi, i, C.byref(x), C.byref(y), C.byref(z))
            assert (x.value, y.value, z.value) == (i + .5, 2. * i, -i)
        v = [C.c_int() for _ in range(5)]
        assert G.get_line(h, ti, 1, *[C.byref(a) for a in v])
        assert [a.value for a in v] == [2, 1, 0, 1, 12]
        n, idx = C.c_int(), C.c_int()
        assert L.GmlGetMeshInfo(h, G.GmlTetrahedra, C.byref(n), C.byref(idx)) and (n.value, idx.value) == (2, ti)
        assert not L.GmlGetMeshInfo(h, G.GmlHexahedra, C.byref(n), C.byref(idx))
        # block get with a 4-float user stride: x, y, z written, the fourth slot left as it was
        crd = np.full((3, 4), 9.0, np.float32)
        ref = np.empty(3, np.int32)
        assert L.GmlGetDataBlock(h, G.GmlVertices, 0, 2, G.ptr(crd), C.c_void_p(crd.ctypes.data + 32),
                                 G.ptr(ref), C.c_void_p(ref.ctypes.data + 8))
        assert np.array_equal(crd[:, 0], [.5, 1.5, 2.5]) and np.array_equal(ref, [0, 7, 14])
        assert np.all(crd[:, 3] == 9.0)


def test_block_set_with_stride_and_shared_ref():
    # the reference infers strides from first/last line pointers; a zero
    # ref stride (same pointer twice) gives every line the same ref
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        vi = L.GmlNewMeshData(h, G.GmlVertices, 4)
        c = np.arange(4 * 5, dtype=np.float32).reshape(4, 5)
        r = np.array([9], np.int32)
        assert L.GmlSetDataBlock(h, G.GmlVertices, 0, 3, G.ptr(c), C.c_void_p(c.ctypes.data + 3 * 20), G.ptr(r), G.ptr(r))
        crd, ref = g.vertices(4)
        assert np.array_equal(crd[:, :3], c[:, :3]) and np.all(ref == 9)
        assert not L.GmlSetDataBlock(h, G.GmlVertices, 2, 2, G.ptr(c), G.ptr(c), None, None)


def test_solution_lines_and_params():
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        L.GmlNewMeshData(h, G.GmlTetrahedra, 5)
        s = g.solution(G.GmlTetrahedra, 3, G.GmlFlt4, "OptCrd")
        line = np.arange(12, dtype=np.float32)
        for i in range(5):
            assert G.set_line(h, s, i, G.ptr(line + i))
        out = np.empty(12, np.float32)
        assert G.get_line(h, s, 0, G.ptr(out)) and np.array_equal(out, line)
        assert G.get_line(h, s, 4, G.ptr(out)) and np.array_equal(out, line + 4)
        p = g.new_parameters(16, "typedef struct { float omega; int pad[3]; } GmlParSct;")
        C.cast(p, C.POINTER(C.c_float))[0] = 0.25
        assert L.GmlUploadParameters(h) == 1
        assert L.GmlGetMemoryUsage(h) >= 5 * 48 + 5 * 16 + 5 * 4 + 16
        assert L.GmlGetMemoryTransfer(h) > 0


def test_error_codes():
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        assert L.GmlNewMeshData(h, 9, 10) == 0
        assert L.GmlNewSolutionData(h, G.GmlVertices, 1, 25, b"x") == 0
        assert L.GmlLaunchKernel(h, 0) == -1 and L.GmlLaunchKernel(h, 57) == -1
        r = C.c_double()
        assert L.GmlReduceVector(h, 0, G.GmlMin, C.byref(r)) == -1
        L.GmlNewMeshData(h, G.GmlVertices, 8)
        f4 = g.solution(G.GmlVertices, 1, G.GmlFlt4, "v4")
        assert L.GmlReduceVector(h, f4, G.GmlMin, C.byref(r)) == -2
        f1 = g.solution(G.GmlVertices, 1, G.GmlFlt, "v1")
        assert L.GmlReduceVector(h, f1, G.GmlMaxRed, C.byref(r)) == -3
        assert L.GmlReduceVector(h, f1, -1, C.byref(r)) == -3
        # no device on a host-only handle: the reduction does not run
        assert L.GmlReduceVector(h, f1, G.GmlSum, C.byref(r)) == -6
        assert L.GmlFreeData(h, f1) == 1 and L.GmlFreeData(h, f1) == 0
        assert L.GmlCheckFP64(h) == 1


def test_data_slot_exhaustion():
    with G.Gml(host_only=True) as g:
        L, h = g.L, g.h
        L.GmlNewMeshData(h, G.GmlVertices, 2)   # 2 slots
        got = [L.GmlNewSolutionData(h, G.GmlVertices, 1, G.GmlFlt, b"s") for _ in range(100)]
        assert sum(1 for x in got if x) == 98 and got[-1] == 0


def test_extract_edges_counts():
    crd, ref, tet = G.cube_mesh(3, 0.0, 1, True)
    with G.Gml(host_only=True) as g:
        g.set_mesh(crd, ref, tet)
        ne = g.L.GmlExtractEdges(g.h)
        # Kuhn split of n^3 cells: 3n(n+1)^2 grid edges + 3n^2(n+1) face
        # diagonals + n^3 cell diagonals
        n = 3
        assert ne == 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
        e = np.empty((ne, 2), np.int32)
        r = np.empty(ne, np.int32)
        assert g.L.GmlGetDataBlock(g.h, G.GmlEdges, 0, ne - 1, G.ptr(e), C.c_void_p(e.ctypes.data + (ne - 1) * 8),
                                   G.ptr(r), C.c_void_p(r.ctypes.data + (ne - 1) * 4))
        assert np.all(e[:, 0] < e[:, 1]) and len({tuple(x) for x in e}) == ne


def test_c_user_program_compiles_and_links():
    """A plain C GMlib program builds against include/gmlib3.h and links
    libgml.so with gcc (running it needs the GPU: tests/test_gpu_parity.py)."""
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "gmlib_amd", "examples", "mesh_quality.c")
    out = os.path.join(tempfile.mkdtemp(), "mesh_quality")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                        "-L", os.path.join(ROOT, "gmlib_amd"), "-lgml",
                        "-Wl,-rpath," + os.path.join(ROOT, "gmlib_amd"), "-o", out],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    assert os.path.exists(out)
