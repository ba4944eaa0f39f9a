"""GmlImportMesh / GmlExportSolution (sources/gmlib3.c:4456-4740) against
an independent reader and writer of the libMeshb formats written here with
struct: binary versions 1-4 in both byte orders and ASCII files of a cube
mesh, the reference's own sample mesh (tests/golden, a 2D libMeshb version
1 file), and solution files read back field by field."""
import ctypes as C
import lzma
import os
import struct

import numpy as np
import pytest

import gmlib_amd as G

HERE = os.path.dirname(os.path.abspath(__file__))
GMF = {"Dimension": 3, "Vertices": 4, "Edges": 5, "Triangles": 6, "Quadrilaterals": 7, "Tetrahedra": 8,
       "Prisms": 9, "Hexahedra": 10, "Pyramids": 49, "End": 54, "SolAtVertices": 62, "SolAtTetrahedra": 66}
KIND = {"Vertices": G.GmlVertices, "Edges": G.GmlEdges, "Triangles": G.GmlTriangles, "Tetrahedra": G.GmlTetrahedra}


def write_meshb(path, ver, dim, blocks, big=False):
    """blocks: [(name, array)] -- vertices (n, dim + 1) as (coords, ref),
    elements (n, nnod + 1) as (1-based nodes, ref)."""
    e = ">" if big else "<"
    I, P = ("q" if ver >= 4 else "i"), ("q" if ver >= 3 else "i")
    R = "d" if ver >= 2 else "f"
    out = bytearray(struct.pack(e + "ii", 1, ver))
    # Dimension
    hdr = struct.calcsize(e + "i" + P)
    out += struct.pack(e + "i" + P + "i", GMF["Dimension"], len(out) + hdr + 4, dim)
    for name, a in blocks:
        body = bytearray()
        for row in a:
            if name == "Vertices":
                body += struct.pack(e + R * dim + I, *[float(x) for x in row[:dim]], int(row[dim]))
            else:
                body += struct.pack(e + I * len(row), *[int(x) for x in row])
        head = struct.calcsize(e + "i" + P + I)
        out += struct.pack(e + "i" + P + I, GMF[name], len(out) + head + len(body), len(a)) + body
    out += struct.pack(e + "i" + P, GMF["End"], 0)
    open(path, "wb").write(out)


def write_mesh_ascii(path, dim, blocks):
    with open(path, "w") as f:
        f.write("MeshVersionFormatted 2\n# a comment line\nDimension\n%d\n\n" % dim)
        f.write("Corners\n2\n1\n2\n\n")   # a keyword the reader skips
        for name, a in blocks:
            f.write("%s\n%d\n" % (name, len(a)))
            for row in a:
                if name == "Vertices":
                    f.write(" ".join("%.9g" % x for x in row[:dim]) + " %d\n" % int(row[dim]))
                else:
                    f.write(" ".join(str(int(x)) for x in row) + "\n")
            f.write("\n")
        f.write("End\n")


def read_meshb(data):
    """Independent parse: {name: (count, raw bytes start)} plus version, dim, order."""
    code = struct.unpack_from("<i", data, 0)[0]
    e = "<" if code == 1 else ">"
    ver = struct.unpack_from(e + "i", data, 4)[0]
    P, I = ("q" if ver >= 3 else "i"), ("q" if ver >= 4 else "i")
    pos, kw, dim = 8, {}, 3
    names = {v: k for k, v in GMF.items()}
    while True:
        k = struct.unpack_from(e + "i", data, pos)[0]
        if k == GMF["End"]:
            break
        nxt = struct.unpack_from(e + P, data, pos + 4)[0]
        at = pos + 4 + struct.calcsize(P)
        if k == GMF["Dimension"]:
            dim = struct.unpack_from(e + "i", data, at)[0]
        elif k in names:
            n = struct.unpack_from(e + I, data, at)[0]
            kw[names[k]] = (n, at + struct.calcsize(I))
        if not nxt:
            break
        pos = nxt
    return ver, dim, e, kw


def mesh_tables(g):
    out = {}
    for name, typ in KIND.items():
        n = C.c_int(0)
        if not g.L.GmlGetMeshInfo(g.h, typ, C.byref(n), None):
            continue
        n = n.value
        w = 4 if typ == G.GmlVertices else {G.GmlEdges: 2, G.GmlTriangles: 3, G.GmlTetrahedra: 4}[typ]
        a = np.zeros((n, w), np.float32 if typ == G.GmlVertices else np.int32)
        r = np.zeros(n, np.int32)
        assert g.L.GmlGetDataBlock(g.h, typ, 0, n - 1, G.ptr(a), C.c_void_p(a.ctypes.data + (n - 1) * w * 4),
                                   G.ptr(r), C.c_void_p(r.ctypes.data + (n - 1) * 4))
        out[name] = (a, r)
    return out


def cube_blocks(n=3):
    crd, ref, tet = G.cube_mesh(n, 0.2, 9)
    rng = np.random.default_rng(1)
    vref = rng.integers(0, 9, len(crd))
    tref = rng.integers(0, 5, len(tet))
    tri = np.array([[0, 1, 2], [3, 4, 5], [len(crd) - 1, 1, 7]])
    triref = np.array([11, 12, 13])
    verts = np.concatenate([crd[:, :3].astype(np.float64), vref[:, None]], 1)
    return (crd, vref, tet, tref, tri, triref,
            [("Vertices", verts), ("Triangles", np.concatenate([tri + 1, triref[:, None]], 1)),
             ("Tetrahedra", np.concatenate([tet + 1, tref[:, None]], 1))])


def check_cube(g, crd, vref, tet, tref, tri, triref):
    t = mesh_tables(g)
    assert np.array_equal(t["Vertices"][0][:, :3], crd[:, :3]) and np.array_equal(t["Vertices"][1], vref)
    assert np.array_equal(t["Tetrahedra"][0], tet) and np.array_equal(t["Tetrahedra"][1], tref)
    assert np.array_equal(t["Triangles"][0], tri) and np.array_equal(t["Triangles"][1], triref)


def imp(g, path, *kwds):
    return g.L.GmlImportMesh(C.c_size_t(g.h), C.c_char_p(path.encode()), *[C.c_int(k) for k in kwds], C.c_int(0))


@pytest.mark.parametrize("ver,big", [(1, False), (2, False), (3, False), (4, False), (1, True), (3, True)])
def test_import_binary_versions(tmp_path, ver, big):
    crd, vref, tet, tref, tri, triref, blocks = cube_blocks()
    path = str(tmp_path / "cube.meshb")
    write_meshb(path, ver, 3, blocks, big)
    with G.Gml(host_only=True) as g:
        # Edges are requested but absent; Corners is not an element keyword
        assert imp(g, path, GMF["Vertices"], GMF["Edges"], GMF["Tetrahedra"], GMF["Triangles"], 13) == 4
        check_cube(g, crd, vref, tet, tref, tri, triref)
        assert not g.L.GmlGetMeshInfo(g.h, G.GmlEdges, None, None)


def test_import_ascii(tmp_path):
    crd, vref, tet, tref, tri, triref, blocks = cube_blocks()
    path = str(tmp_path / "cube.mesh")
    write_mesh_ascii(path, 3, blocks)
    with G.Gml(host_only=True) as g:
        assert imp(g, path, GMF["Vertices"], GMF["Triangles"], GMF["Tetrahedra"]) == 3
        check_cube(g, crd, vref, tet, tref, tri, triref)


def test_import_only_requested_and_missing_file(tmp_path):
    crd, vref, tet, tref, tri, triref, blocks = cube_blocks()
    path = str(tmp_path / "cube.meshb")
    write_meshb(path, 1, 3, blocks)
    with G.Gml(host_only=True) as g:
        assert imp(g, path, GMF["Vertices"]) == 1
        assert g.L.GmlGetMeshInfo(g.h, G.GmlVertices, None, None)
        assert not g.L.GmlGetMeshInfo(g.h, G.GmlTetrahedra, None, None)
        assert imp(g, str(tmp_path / "none.meshb"), GMF["Vertices"]) == 0


def test_import_reference_sample_mesh(tmp_path):
    data = lzma.open(os.path.join(HERE, "golden", "naca012_2D_triangles.meshb.xz")).read()
    ver, dim, e, kw = read_meshb(data)
    assert (ver, dim) == (1, 2) and {k: v[0] for k, v in kw.items()} == {"Vertices": 36624, "Edges": 465,
                                                                        "Triangles": 72783}
    path = str(tmp_path / "naca.meshb")
    open(path, "wb").write(data)
    with G.Gml(host_only=True) as g:
        assert imp(g, path, GMF["Vertices"], GMF["Edges"], GMF["Triangles"]) == 3
        t = mesh_tables(g)
    n, at = kw["Vertices"]
    v = np.frombuffer(data, np.dtype([("x", e + "f4"), ("y", e + "f4"), ("r", e + "i4")]), n, at)
    assert np.array_equal(t["Vertices"][0][:, 0], v["x"]) and np.array_equal(t["Vertices"][0][:, 1], v["y"])
    assert (t["Vertices"][0][:, 2] == 0).all() and np.array_equal(t["Vertices"][1], v["r"])
    for name, nn in (("Edges", 2), ("Triangles", 3)):
        n, at = kw[name]
        a = np.frombuffer(data, np.dtype(e + "i4"), n * (nn + 1), at).reshape(n, nn + 1)
        assert np.array_equal(t[name][0], a[:, :nn] - 1) and np.array_equal(t[name][1], a[:, nn])
        assert t[name][0].min() >= 0 and t[name][0].max() < 36624


def read_sol_binary(data):
    code, ver = struct.unpack_from("<ii", data, 0)
    assert code == 1 and ver == 1
    pos, out = 8, {}
    while True:
        k, nxt = struct.unpack_from("<ii", data, pos)
        if k == GMF["End"]:
            break
        if k == GMF["Dimension"]:
            assert struct.unpack_from("<i", data, pos + 8)[0] == 3
        else:
            n, nt = struct.unpack_from("<ii", data, pos + 8)
            typ = struct.unpack_from("<%di" % nt, data, pos + 16)
            vals = np.frombuffer(data, "<f4", n * nt, pos + 16 + 4 * nt).reshape(n, nt)
            out[k] = (typ, vals)
            assert nxt == pos + 16 + 4 * nt + 4 * n * nt
        pos = nxt
    return out


def test_export_solution_binary_and_ascii(tmp_path):
    crd, ref, tet = G.cube_mesh(2, 0.2, 4)
    T, V = len(tet), len(crd)
    with G.Gml(host_only=True) as g:
        g.set_mesh(crd, ref, tet)
        q = np.random.default_rng(0).random(T).astype(np.float32)
        v4 = np.random.default_rng(1).random((V, 4)).astype(np.float32)
        h2 = np.random.default_rng(2).random((T, 2)).astype(np.float64)
        qi = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        vi = g.solution(G.GmlVertices, 1, G.GmlFlt4, "vec")
        hi = g.solution(G.GmlTetrahedra, 1, G.GmlDbl2, "grd")
        for idx, a in ((qi, q), (vi, v4), (hi, h2)):
            for i in range(len(a)):
                G.set_line(g.h, idx, i, C.c_void_p(a[i:i + 1].ctypes.data))
        binp, ascp = str(tmp_path / "out.solb"), str(tmp_path / "out.sol")
        for p in (binp, ascp):
            assert g.L.GmlExportSolution(C.c_size_t(g.h), C.c_char_p(p.encode()), C.c_int(qi), C.c_int(vi),
                                         C.c_int(hi), C.c_int(0)) == 2
    sol = read_sol_binary(open(binp, "rb").read())
    typ, vals = sol[GMF["SolAtTetrahedra"]]
    assert typ == (1, 1, 1) and vals.shape == (T, 3)
    assert np.array_equal(vals[:, 0], q) and np.array_equal(vals[:, 1:], h2.astype(np.float32))
    typ, vals = sol[GMF["SolAtVertices"]]
    assert typ == (1, 1, 1, 1) and np.array_equal(vals, v4)
    # ASCII: same fields, then the field names
    txt = open(ascp).read().split()
    i = txt.index("SolAtTetrahedra")
    assert txt[i + 1:i + 6] == [str(T), "3", "1", "1", "1"]
    a = np.array(txt[i + 6:i + 6 + 3 * T], np.float64).reshape(T, 3)
    assert np.allclose(a[:, 0], q, rtol=1e-7, atol=0) and np.allclose(a[:, 1:], h2, rtol=1e-7, atol=0)
    j = txt.index("ReferenceStrings")
    assert txt[j + 1] == "7" and txt[j + 2:j + 6] == ["66", "1", "qal", "1"]
    assert "vec.x" in txt and "grd.y" in txt


SAMPLES = "/root/reference/sample_meshes"


@pytest.mark.skipif(not os.path.isdir(SAMPLES), reason="reference tree not present")
@pytest.mark.parametrize("name,kinds", [("edges", ("Edges",)), ("triangles", ("Triangles",))])
def test_import_reference_3d_sample_meshes(tmp_path, name, kinds):
    # the reference's other sample meshes (version 1, dimension 3), read
    # where they lie and compared field by field with the independent parse
    data = lzma.open(os.path.join(SAMPLES, name + ".meshb.xz")).read()
    ver, dim, e, kw = read_meshb(data)
    assert (ver, dim) == (1, 3) and set(kw) == {"Vertices"} | set(kinds)
    path = str(tmp_path / (name + ".meshb"))
    open(path, "wb").write(data)
    with G.Gml(host_only=True) as g:
        assert imp(g, path, GMF["Vertices"], *[GMF[k] for k in kinds]) == 1 + len(kinds)
        t = mesh_tables(g)
    n, at = kw["Vertices"]
    v = np.frombuffer(data, np.dtype([("x", e + "f4"), ("y", e + "f4"), ("z", e + "f4"), ("r", e + "i4")]), n, at)
    assert np.array_equal(t["Vertices"][0][:, 0], v["x"]) and np.array_equal(t["Vertices"][0][:, 2], v["z"])
    assert (t["Vertices"][0][:, 3] == 0).all() and np.array_equal(t["Vertices"][1], v["r"])
    for k in kinds:
        nn = {"Edges": 2, "Triangles": 3}[k]
        m, at = kw[k]
        a = np.frombuffer(data, np.dtype(e + "i4"), m * (nn + 1), at).reshape(m, nn + 1)
        assert np.array_equal(t[k][0], a[:, :nn] - 1) and np.array_equal(t[k][1], a[:, nn])
        assert t[k][0].min() >= 0 and t[k][0].max() < n


@pytest.mark.skipif(not os.path.isdir(SAMPLES), reason="reference tree not present")
def test_extract_edges_of_reference_surface_mesh(tmp_path):
    # GmlExtractEdges on the reference's triangle sample (a host-only handle:
    # the ordered-map form) against the oracle's restatement of the
    # reference hash: the same edges in the same order; every edge of a
    # closed manifold surface is shared by two triangles (E = 3T / 2)
    from tests import oracle_lib as O
    data = lzma.open(os.path.join(SAMPLES, "triangles.meshb.xz")).read()
    path = str(tmp_path / "triangles.meshb")
    open(path, "wb").write(data)
    with G.Gml(host_only=True) as g:
        assert imp(g, path, GMF["Vertices"], GMF["Triangles"]) == 2
        t = mesh_tables(g)
        ne = g.L.GmlExtractEdges(g.h)
        assert ne > 0
        tab = mesh_tables(g)["Edges"]
    tri = t["Triangles"][0]
    want, wref = O.extract_items(2, [(G.GmlTriangles, tri)])
    assert np.array_equal(tab[0], want) and ne == len(want) and (tab[1] == 0).all()
    # the same edge set from a plain numpy count (each triangle side once)
    V = len(t["Vertices"][0])
    sides = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1).astype(np.int64)
    keys, cnt = np.unique(sides[:, 0] * V + sides[:, 1], return_counts=True)
    assert len(keys) == ne
    assert np.array_equal(np.sort(want[:, 0].astype(np.int64) * V + want[:, 1]), keys)
    assert (cnt == 2).all() and 2 * ne == 3 * len(tri)       # a closed surface


def _bin(ver, words):
    """a binary file from (format, value) pairs after the 1 / version header"""
    out = bytearray(struct.pack("<ii", 1, ver))
    for f, v in words:
        out += struct.pack("<" + f, v)
    return bytes(out)


@pytest.mark.timeout(30)
def test_import_rejects_malformed_binary_files(tmp_path):
    # the reader does not trust the file: negative or oversized counts, a
    # next-keyword position that does not advance (it would loop), past the
    # end of the file, or an impossible dimension give 0 and no table
    V = GMF["Vertices"]
    cases = {
        "negative_count": [("i", GMF["Dimension"]), ("i", 20), ("i", 3), ("i", V), ("i", 32), ("i", -5)],
        "count_beyond_file": [("i", GMF["Dimension"]), ("i", 20), ("i", 3), ("i", V), ("i", 32), ("i", 1 << 30)],
        "position_loops": [("i", GMF["Dimension"]), ("i", 8), ("i", 3)],
        "position_past_end": [("i", GMF["Dimension"]), ("i", 1 << 20), ("i", 3)],
        "bad_dimension": [("i", GMF["Dimension"]), ("i", 20), ("i", 1000000)],
    }
    for name, words in cases.items():
        path = str(tmp_path / (name + ".meshb"))
        open(path, "wb").write(_bin(1, words))
        with G.Gml(host_only=True) as g:
            assert imp(g, path, GMF["Vertices"]) == 0, name
            assert not g.L.GmlGetMeshInfo(g.h, G.GmlVertices, None, None), name


def test_import_rejects_out_of_range_nodes(tmp_path):
    # a tet referencing vertex 9 of a file with 4 vertices: the tets are not
    # installed (a generated gather would read outside the vertex table)
    verts = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], float)
    tets = np.array([[1, 2, 3, 9, 0]])
    for ver in (1, 4):
        path = str(tmp_path / ("bad%d.meshb" % ver))
        write_meshb(path, ver, 3, [("Vertices", verts), ("Tetrahedra", tets)])
        with G.Gml(host_only=True) as g:
            imp(g, path, GMF["Vertices"], GMF["Tetrahedra"])
            assert g.L.GmlGetMeshInfo(g.h, G.GmlVertices, None, None)
            assert not g.L.GmlGetMeshInfo(g.h, G.GmlTetrahedra, None, None)
    path = str(tmp_path / "bad.mesh")
    write_mesh_ascii(path, 3, [("Vertices", verts), ("Tetrahedra", tets)])
    with G.Gml(host_only=True) as g:
        imp(g, path, GMF["Vertices"], GMF["Tetrahedra"])
        assert not g.L.GmlGetMeshInfo(g.h, G.GmlTetrahedra, None, None)
    # an ASCII count far beyond the file
    path = str(tmp_path / "huge.mesh")
    open(path, "w").write("MeshVersionFormatted 2\nDimension\n3\nVertices\n99999999999\n0 0 0 1\nEnd\n")
    with G.Gml(host_only=True) as g:
        assert imp(g, path, GMF["Vertices"]) == 0
