"""GPU tests of the serving loop (GmlxServeStart / Submit / Wait).

The same jobs are run alone (plain path: a fresh handle per job, one job at
a time) and in a mix (4 slots, all submitted at once, an HBM budget small
enough to force evictions and re-loads).  Every result must be the same in
both: quality and smoothing results (min and mean quality) bit-identical,
CG results (FP64 relative residual, iteration count) identical -- every
kernel of the path is deterministic (fixed-order reductions), and a mesh
serves one job at a time from its original coordinates."""
import ctypes as C

import numpy as np
import pytest

import gmlib_amd as G

pytestmark = pytest.mark.gpu


def job_list(count, seed=3):
    rng = np.random.default_rng(seed)
    meshes = [(3, 101), (4, 102), (5, 103), (6, 104), (8, 105), (4, 106)]
    jobs = []
    for _ in range(count):
        n, s = meshes[rng.integers(len(meshes))]
        t = int(rng.choice([G.GmlxJobQuality, G.GmlxJobSmooth, G.GmlxJobCG], p=[0.5, 0.3, 0.2]))
        param = {G.GmlxJobQuality: 0.0, G.GmlxJobSmooth: 5.0, G.GmlxJobCG: 1e-8}[t]
        jobs.append((t, n, s, param))
    return jobs


def run(srv, jobs):
    L = G.lib()
    ids = [L.GmlxServeSubmit(srv, t, n, s, p) for t, n, s, p in jobs]
    assert all(i > 0 for i in ids)
    out = []
    for i in ids:
        res = (C.c_double * 4)()
        tim = (C.c_double * 4)()
        rc = L.GmlxServeWait(srv, i, res, tim)
        out.append((rc, tuple(res), tuple(tim)))
    return out


def test_serve_same_alone_and_in_the_mix():
    L = G.lib()
    jobs = job_list(30)
    plain = L.GmlxServeStart(0, 1.0, 1, 0)
    assert plain
    alone = run(plain, jobs)
    info = (C.c_double * 8)()
    assert L.GmlxServeInfo(plain, info) and info[5] == 1        # one job at a time
    L.GmlxServeStop(plain)

    srv = L.GmlxServeStart(0, 0.008, 4, 1)                      # 8 MB: a few small meshes at a time
    assert srv
    mix = run(srv, jobs)
    assert L.GmlxServeInfo(srv, info)
    resident, used, loads, evictions, done, maxrun, budget, peak = list(info)
    L.GmlxServeStop(srv)
    print("mix: loads %d, evictions %d, most jobs at once %d, peak %.1f MB of %.1f MB"
          % (loads, evictions, maxrun, peak / 1e6, budget / 1e6))
    assert done == len(jobs) and maxrun > 1 and evictions > 0
    for (t, n, s, p), a, m in zip(jobs, alone, mix):
        assert a[0] == 1 and m[0] == 1, (t, n, s, a[0], m[0])
        assert a[1][:2] == m[1][:2], "job %s: alone %s, in the mix %s" % ((t, n, s), a[1][:2], m[1][:2])
        if t == G.GmlxJobCG:
            assert a[1][0] <= 1e-8 and a[1][1] > 0
        else:
            assert -1.0 <= a[1][0] <= a[1][1] <= 1.0
        assert m[2][0] <= m[2][1] <= m[2][2]                     # submitted <= admitted <= done


def test_serve_smoothing_improves_and_repeats():
    L = G.lib()
    srv = L.GmlxServeStart(0, 1.0, 2, 1)
    out = run(srv, [(G.GmlxJobQuality, 6, 7, 0.0), (G.GmlxJobSmooth, 6, 7, 20.0),
                    (G.GmlxJobQuality, 6, 7, 0.0), (G.GmlxJobSmooth, 6, 7, 20.0)])
    L.GmlxServeStop(srv)
    q0, s0, q1, s1 = (o[1] for o in out)
    assert q0 == q1                     # quality after a smoothing job: original coordinates again
    assert s0 == s1
    assert s0[1] > q0[1]                # smoothing raises the mean quality


def test_serve_over_budget_job_is_refused():
    L = G.lib()
    srv = L.GmlxServeStart(0, 0.002, 1, 1)                      # 2 MB
    out = run(srv, [(G.GmlxJobQuality, 30, 1, 0.0), (G.GmlxJobQuality, 2, 1, 0.0)])
    L.GmlxServeStop(srv)
    assert out[0][0] == -7 and out[1][0] == 1


def test_serve_registered_mesh_matches_cube_job():
    """A registered copy of cube mesh (5, 77) serves the same quality and
    smoothing results as the cube job that names it; CG solves converge."""
    L = G.lib()
    crd, ref, tet = G.cube_mesh(5, 0.2, 77)
    srv = L.GmlxServeStart(0, 0.05, 2, 1)
    key = L.GmlxServeAddMesh(srv, len(crd), len(tet), G.ptr(crd), G.ptr(ref), G.ptr(tet))
    assert key > 0
    bad = tet.copy()
    bad[0, 0] = len(crd)
    assert L.GmlxServeAddMesh(srv, len(crd), len(tet), G.ptr(crd), G.ptr(ref), G.ptr(bad)) == -1
    assert L.GmlxServeSubmitMesh(srv, G.GmlxJobQuality, key + 100, 0.0) == -1
    ids = [L.GmlxServeSubmitMesh(srv, G.GmlxJobQuality, key, 0.0), L.GmlxServeSubmitMesh(srv, G.GmlxJobSmooth, key, 10.0),
           L.GmlxServeSubmitMesh(srv, G.GmlxJobCG, key, 1e-8),
           L.GmlxServeSubmit(srv, G.GmlxJobQuality, 5, 77, 0.0), L.GmlxServeSubmit(srv, G.GmlxJobSmooth, 5, 77, 10.0)]
    out = []
    for i in ids:
        res, tim = (C.c_double * 4)(), (C.c_double * 4)()
        assert L.GmlxServeWait(srv, i, res, tim) == 1
        out.append(tuple(res))
    L.GmlxServeStop(srv)
    assert out[0][:2] == out[3][:2] and out[1][:2] == out[4][:2]
    assert out[2][0] <= 1e-8 and out[2][1] > 0


def _wait_thread(L, srv, jid, out, tag):
    import threading

    def w():
        res, tim = (C.c_double * 4)(), (C.c_double * 4)()
        out[tag] = (L.GmlxServeWait(srv, jid, res, tim), tuple(res))
    t = threading.Thread(target=w)
    t.start()
    return t


def test_serve_two_waiters_one_job_then_id_retired():
    L = G.lib()
    srv = L.GmlxServeStart(0, 1.0, 1, 1)
    jid = L.GmlxServeSubmit(srv, G.GmlxJobSmooth, 10, 3, 20.0)
    out = {}
    ts = [_wait_thread(L, srv, jid, out, k) for k in ("a", "b")]
    for t in ts:
        t.join(60)
    assert out["a"] == out["b"] and out["a"][0] == 1
    res, tim = (C.c_double * 4)(), (C.c_double * 4)()
    assert L.GmlxServeWait(srv, jid, res, tim) == -1            # retired
    L.GmlxServeStop(srv)


def test_serve_stop_ends_pending_jobs_and_wakes_waiters():
    L = G.lib()
    import time
    srv = L.GmlxServeStart(0, 4.0, 1, 1)
    # one slot: a smoothing job of 100000 sweeps (at least half a second of
    # launches) holds it while the quality jobs behind it are still pending
    # when the server stops; the waiters are blocked in GmlxServeWait by then
    first = L.GmlxServeSubmit(srv, G.GmlxJobSmooth, 24, 9, 100000.0)
    rest = [L.GmlxServeSubmit(srv, G.GmlxJobQuality, 6, 9, 0.0) for _ in range(8)]
    out = {}
    ts = [_wait_thread(L, srv, j, out, j) for j in [first] + rest]
    time.sleep(0.2)
    L.GmlxServeStop(srv)
    for t in ts:
        t.join(60)
        assert not t.is_alive()
    assert out[first][0] == 1
    rcs = [out[j][0] for j in rest]
    assert set(rcs) <= {1, -8} and rcs[-1] == -8, rcs


def test_serve_remove_mesh():
    L = G.lib()
    crd, ref, tet = G.cube_mesh(4, 0.2, 5)
    srv = L.GmlxServeStart(0, 0.05, 2, 1)
    key = L.GmlxServeAddMesh(srv, len(crd), len(tet), G.ptr(crd), G.ptr(ref), G.ptr(tet))
    jid = L.GmlxServeSubmitMesh(srv, G.GmlxJobQuality, key, 0.0)
    res, tim = (C.c_double * 4)(), (C.c_double * 4)()
    assert L.GmlxServeWait(srv, jid, res, tim) == 1
    info = (C.c_double * 8)()
    assert L.GmlxServeInfo(srv, info) and info[0] == 1          # resident
    assert L.GmlxServeRemoveMesh(srv, key) == 1
    assert L.GmlxServeInfo(srv, info) and info[0] == 0 and info[1] == 0
    assert L.GmlxServeRemoveMesh(srv, key) == 0
    assert L.GmlxServeSubmitMesh(srv, G.GmlxJobQuality, key, 0.0) == -1
    L.GmlxServeStop(srv)
