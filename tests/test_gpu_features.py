"""GPU parity of API features a user body relies on beyond gathers
(sources/gmlib3.c): GmlManual data (the body indexes the data table itself,
:2683, 2771, 2917) with OpenCL atomic functions, a user toolkit
(GmlIncludeUserToolkit, :4445-4453) whose OpenCL-C functions the body calls,
compiler options (GmlSetCompilerOptions, :2993), and the work-item
functions (get_global_id ...), which name the row as in the reference's
generated counter.  Expected values come from numpy on the same inputs."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G

pytestmark = pytest.mark.gpu


def test_manual_data_with_atomics():
    # per vertex: the number of tets holding it (atomic_inc), the largest and
    # smallest of their indices (atomic_max / atomic_min) and the sum of
    # their local node numbers (atomic_add), through the tets' own table
    crd, ref, tet = G.cube_mesh(12, 0.2, 4)
    V, T = len(crd), len(tet)
    body = """{
       for(int j = 0; j < 4; j++)
       {
          atomic_inc(&cntTab[TetVer[j]]);
          atomic_max(&hiTab[TetVer[j]], TetIdx);
          atomic_min(&loTab[TetVer[j]], TetIdx);
          atomic_add(&locTab[TetVer[j]], j);
       }
    }"""
    RWM = G.GmlReadMode | G.GmlWriteMode | G.GmlManual
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        d = {n: g.solution(G.GmlVertices, 1, G.GmlInt, n) for n in ("cnt", "hi", "lo", "loc")}
        g.upload(d["cnt"], np.zeros(V, np.int32))
        g.upload(d["hi"], np.full(V, -1, np.int32))
        g.upload(d["lo"], np.full(V, T, np.int32))
        g.upload(d["loc"], np.zeros(V, np.int32))
        k = g.compile(body, "vertex_counts", G.GmlTetrahedra, [(d[n], RWM, 0) for n in ("cnt", "hi", "lo", "loc")])
        g.launch(k)
        got = {n: g.download_raw(d[n], np.int32, (V,)) for n in d}
        assert g.L.GmlxGetKernelForm(g.h, k) == 1          # atomics: one row per work-item only
    flat, idx = tet.ravel(), np.repeat(np.arange(T, dtype=np.int32), 4)
    hi = np.full(V, -1, np.int32)
    lo = np.full(V, T, np.int32)
    np.maximum.at(hi, flat, idx)
    np.minimum.at(lo, flat, idx)
    loc = np.zeros(V, np.int64)
    np.add.at(loc, flat, np.tile(np.arange(4), T))
    assert np.array_equal(got["cnt"], np.bincount(flat, minlength=V))
    assert np.array_equal(got["hi"], hi) and np.array_equal(got["lo"], lo)
    assert np.array_equal(got["loc"], loc)


def test_user_toolkit_and_compiler_options():
    crd, ref, tet = G.cube_mesh(8, 0.2, 6)
    T = len(tet)
    toolkit = C.c_char_p(b"float twice(float x) { return 2.f * x; }\n"
                         b"float4 mid(float4 a, float4 b) { return (a + b) * .5f; }\n")
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        g.L.GmlIncludeUserToolkit(g.h, toolkit)
        g.L.GmlSetCompilerOptions(g.h, C.c_char_p(b"-DGMLX_SCALE=3 -cl-mad-enable"))
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile("{ qal = GMLX_SCALE * twice(mid(VerCrd[0], VerCrd[1]).x); }", "toolkit_body",
                      G.GmlTetrahedra, [(vi, G.GmlReadMode, 0), (q, G.GmlWriteMode, 0)])
        g.launch(k)
        got = g.download_raw(q, np.float32, (T,))
    c = crd.astype(np.float32)
    m = (c[tet[:, 0], 0] + c[tet[:, 1], 0]) * np.float32(0.5)
    want = np.float32(3) * (np.float32(2) * m)
    assert np.array_equal(got, want)


def test_work_item_functions_name_the_row():
    crd, ref, tet = G.cube_mesh(10, 0.2, 8)
    T = len(tet)
    body = """{
       qal = (float)get_global_id(0);
       grp = (get_group_id(0) * get_local_size(0) + get_local_id(0) == get_global_id(0)
              && get_global_size(0) == get_num_groups(0) * get_local_size(0)) ? 1 : 0;
    }"""
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        ok = g.solution(G.GmlTetrahedra, 1, G.GmlInt, "grp")
        k = g.compile(body, "row_ids", G.GmlTetrahedra, [(q, G.GmlWriteMode, 0), (ok, G.GmlWriteMode, 0)])
        assert not g.L.GmlxSetKernelShape(g.h, k, 128, 1, 1)     # no remap for such a body
        for _ in range(6):                                        # through the tuner's shapes
            g.launch(k)
            assert np.array_equal(g.download_raw(q, np.float32, (T,)), np.arange(T, dtype=np.float32))
            assert g.download_raw(ok, np.int32, (T,)).all()
        a, b = C.c_int(), C.c_int()
        g.L.GmlxGetKernelShape(g.h, k, C.byref(a), C.byref(b))
        assert a.value in (64, 128, 256) and b.value == 0


def test_transverse_reads_through_neighbours_and_user_links():
    # a tet field read through the tet -> tet neighbour link (GmlSetNeighbours)
    # and through a link the user fills (GmlNewLinkData); entry 0 means none
    # and reads the null value (sources/gmlib3.c:2788-2790, 2838-2842)
    crd, ref, tet = G.cube_mesh(6, 0.2, 9)
    T = len(tet)
    rng = np.random.default_rng(11)
    val0 = rng.uniform(1.0, 2.0, T).astype(np.float32)
    pair = ((np.arange(T) * 7919) % T).astype(np.int32)
    pair[::5] = 0
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        ngb = g.L.GmlSetNeighbours(g.h, G.GmlTetrahedra)
        assert ngb
        v = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "val")
        g.upload(v, val0)
        s = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "nsum")
        k = g.compile("{ nsum = val[0] + val[1] + val[2] + val[3]; }", "neighbour_sum", G.GmlTetrahedra,
                      [(v, G.GmlReadMode, ngb), (s, G.GmlWriteMode, 0)])
        g.launch(k)
        nsum = g.download_raw(s, np.float32, (T,))
        table = g.download_raw(ngb, np.int32, (T, 4))
        lnk = g.L.GmlNewLinkData(g.h, G.GmlTetrahedra, G.GmlTetrahedra, 1, g.cstr("pair"))
        assert lnk
        g.upload(lnk, pair)
        o = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "other")
        k2 = g.compile("{ other = val; }", "pair_read", G.GmlTetrahedra, [(v, G.GmlReadMode, lnk), (o, G.GmlWriteMode, 0)])
        g.launch(k2)
        other = g.download_raw(o, np.float32, (T,))
    want = np.zeros(T, np.float32)
    for c in range(4):
        want = want + np.where(table[:, c] != 0, val0[table[:, c]], np.float32(0))
    assert (table != 0).any() and (table == 0).any()
    assert np.array_equal(nsum, want)
    assert np.array_equal(other, np.where(pair != 0, val0[pair], np.float32(0)))


def test_parameters_round_trip_and_counters():
    # a body reads a parameter and one row writes another; the host reads it
    # back with GmlDownloadParameters; the run-time and byte counters move
    crd, ref, tet = G.cube_mesh(6, 0.2, 10)
    T = len(tet)
    src = "typedef struct { float scale; int seen; int pad[2]; } GmlParSct;"
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        par = C.cast(g.new_parameters(16, src), C.POINTER(C.c_float))
        par[0] = 2.5
        C.cast(par, C.POINTER(C.c_int))[1] = 0
        assert g.L.GmlUploadParameters(g.h) == 1
        q = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "qal")
        k = g.compile("{ qal = GmlPar->scale * (float)TetIdx; if(TetIdx == 3) GmlPar->seen = 42; }", "par_rw",
                      G.GmlTetrahedra, [(q, G.GmlWriteMode, 0)])
        t0 = g.L.GmlGetWallClock()
        g.launch(k)
        got = g.download_raw(q, np.float32, (T,))
        assert g.L.GmlDownloadParameters(g.h) == 1
        assert C.cast(par, C.POINTER(C.c_int))[1] == 42 and par[0] == 2.5
        assert g.reduce(q, G.GmlMax) == float(np.float32(2.5) * np.float32(T - 1))
        assert g.L.GmlGetKernelRunTime(g.h, k) > 0 and g.L.GmlGetReduceRunTime(g.h, G.GmlMax) > 0
        assert g.L.GmlGetWallClock() >= t0
        assert g.L.GmlGetMemoryUsage(g.h) > 0 and g.L.GmlGetMemoryTransfer(g.h) > 0
    assert np.array_equal(got, np.float32(2.5) * np.arange(T, dtype=np.float32))


@pytest.mark.parametrize("typ,dtype,width", [(G.GmlByt, np.int8, 1), (G.GmlByt4, np.int8, 4), (G.GmlInt2, np.int32, 2),
                                             (G.GmlFlt8, np.float32, 8), (G.GmlDbl4, np.float64, 4),
                                             (G.GmlDbl16, np.float64, 16)])
def test_uplink_reads_of_every_width_staged_and_direct(typ, dtype, width):
    # tet data of 1 to 128 bytes read through the vertex -> tet ball, with the
    # LDS staging on (several items per thread and pass, by item size) and
    # off: every vertex gets its last ball entry's value (the smallest tet of
    # its ball), compared exactly
    crd, ref, tet = G.cube_mesh(9, 0.2, 12)
    V, T = len(crd), len(tet)
    rng = np.random.default_rng(width)
    if dtype == np.int8:
        x0 = rng.integers(-128, 128, size=(T, width)).astype(np.int8)
    else:
        x0 = rng.normal(size=(T, width)).astype(dtype)
    body = "{ y = x[0]; for(int i = 1; i < xDeg; i++) y = x[i]; }"
    outs = []
    for staged in (1, 0):
        with G.Gml() as g:
            G.lib().GmlxSetStaging(g.h, staged)
            vi, ti = g.set_mesh(crd, ref, tet)
            x = g.solution(G.GmlTetrahedra, 1, typ, "x")
            y = g.solution(G.GmlVertices, 1, typ, "y")
            g.upload(x, x0)
            k = g.compile(body, "last_of_ball", G.GmlVertices, [(x, G.GmlReadMode, 0), (y, G.GmlWriteMode, 0)])
            for _ in range(4):
                g.launch(k)
            outs.append(g.download_raw(y, dtype, (V, width)))
            deg, ball = g.uplink_table(G.GmlVertices, G.GmlTetrahedra, V, 32)
    want = x0[ball[np.arange(V), deg - 1] >> 4]
    assert np.array_equal(outs[0], want) and np.array_equal(outs[1], want)


@pytest.mark.parametrize("fields", [[(G.GmlByt2, np.int8, 2)], [(G.GmlDbl16, np.float64, 16)],
                                    [(G.GmlDbl16, np.float64, 16)] * 3])
def test_downlink_reads_of_every_width_in_every_form(fields):
    # vertex data of 2 to 128 bytes (three 128-byte fields: more than the
    # element-tile form's LDS takes) read through the tet -> vertex downlink;
    # every launch of the tuner's shapes gives the sum of the tet's corner 0
    # and corner 3 values, exactly
    crd, ref, tet = G.cube_mesh(9, 0.2, 13)
    V, T = len(crd), len(tet)
    rng = np.random.default_rng(len(fields))
    xs = [(rng.integers(-60, 60, size=(V, w)).astype(dt) if dt == np.int8 else
           rng.normal(size=(V, w)).astype(dt)) for _, dt, w in fields]
    body = "{ " + " ".join("z%d = x%d[0] + x%d[3];" % (i, i, i) for i in range(len(fields))) + " }"
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        args, zs = [], []
        for i, (typ, dt, w) in enumerate(fields):
            x = g.solution(G.GmlVertices, 1, typ, "x%d" % i)
            g.upload(x, xs[i])
            z = g.solution(G.GmlTetrahedra, 1, typ, "z%d" % i)
            args += [(x, G.GmlReadMode, 0)]
            zs.append(z)
        args += [(z, G.GmlWriteMode, 0) for z in zs]
        k = g.compile(body, "corner_sum", G.GmlTetrahedra, args)
        for _ in range(20):
            g.launch(k)
            for i, (typ, dt, w) in enumerate(fields):
                got = g.download_raw(zs[i], dt, (T, w))
                assert np.array_equal(got, (xs[i][tet[:, 0]] + xs[i][tet[:, 3]]).astype(dt))
        tiled = bool(g.L.GmlxGetKernelTileSource(g.h, k))
    assert tiled == (sum(f[0] != G.GmlByt2 and 128 or 2 for f in fields) * 256 <= 64 * 1024)


def test_downlink_and_neighbour_reads_in_one_kernel():
    # a tet kernel reading its vertices' coordinates (downlink: the element-
    # tile form stages them) and a tet field through the neighbour link
    # (read directly): the same values in every form the tuner tries
    crd, ref, tet = G.cube_mesh(8, 0.2, 14)
    T = len(tet)
    val0 = np.random.default_rng(2).uniform(1, 2, T).astype(np.float32)
    with G.Gml() as g:
        vi, ti = g.set_mesh(crd, ref, tet)
        ngb = g.L.GmlSetNeighbours(g.h, G.GmlTetrahedra)
        v = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "val")
        g.upload(v, val0)
        o = g.solution(G.GmlTetrahedra, 1, G.GmlFlt, "mix")
        k = g.compile("{ mix = VerCrd[0].x + VerCrd[2].y + val[1]; }", "mixed_links", G.GmlTetrahedra,
                      [(vi, G.GmlReadMode, 0), (v, G.GmlReadMode, ngb), (o, G.GmlWriteMode, 0)])
        table = g.download_raw(ngb, np.int32, (T, 4))
        c = crd.astype(np.float32)
        want = (c[tet[:, 0], 0] + c[tet[:, 2], 1]) + np.where(table[:, 1] != 0, val0[table[:, 1]], np.float32(0))
        forms = set()
        for _ in range(20):
            g.launch(k)
            assert np.array_equal(g.download_raw(o, np.float32, (T,)), want)
            forms.add(g.L.GmlxGetKernelForm(g.h, k))
    assert forms - {0}
