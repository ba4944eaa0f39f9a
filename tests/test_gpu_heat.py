"""The cell-centered heat solver (gmlib_amd/heat.py, the reference's
HeatCellCentered application) on the GPU against the numpy FP64
restatement of its scheme: temperatures and residuals after a fixed number
of steps, and the converged solution and step count."""
import numpy as np
import pytest

import gmlib_amd as G
from gmlib_amd import heat as H
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def reference_tables(crd, tet, tri, side):
    return O.extract_items(3, [(G.GmlTetrahedra, tet)], tri, side)


@pytest.mark.parametrize("n,steps", [(3, 25), (6, 40)])
def test_heat_steps_match_fp64(n, steps):
    crd, vref, tet, tri, side = H.heat_cube(n, 0.2, 7)
    s = H.HeatSolver(crd, vref, tet, tri, side)
    try:
        res = s.iterate(steps)
        u = s.temperature()
        dt = s.par.dt
        assert s.nface == (4 * len(tet) + 12 * n * n) // 2
    finally:
        s.close()
    full, fref = reference_tables(crd, tet, tri, side)
    u64, res64, dt64 = O.heat_iterations(crd, tet, full, fref, H.DEFAULT_BC, steps, dt=float(dt))
    assert abs(dt - dt64) <= 1e-6 * dt64
    print("n=%d: max |T - T64| = %.2e after %d steps" % (n, np.abs(u - u64).max(), steps))
    assert np.abs(u - u64).max() < 1e-5 * 300 * steps ** 0.5
    assert np.allclose(res, res64, rtol=1e-3, atol=0)


def test_heat_converges_like_fp64():
    n = 3
    crd, vref, tet, tri, side = H.heat_cube(n, 0.2, 3)
    s = H.HeatSolver(crd, vref, tet, tri, side)
    try:
        it = s.solve(drop=1e-6, chunk=50)
        u = s.temperature()
        dt = float(s.par.dt)
    finally:
        s.close()
    full, fref = reference_tables(crd, tet, tri, side)
    u64, res64, _ = O.heat_iterations(crd, tet, full, fref, H.DEFAULT_BC, it, dt=dt)
    first = next(i for i, r in enumerate(res64) if r / res64[0] <= 1e-6) + 1
    print("converged in %d steps (FP64: residual drop reached at step %d)" % (it, first))
    assert first <= it <= first + 50
    assert np.abs(u - u64).max() < 0.05
