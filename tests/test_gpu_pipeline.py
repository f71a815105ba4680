"""Pipelined sweeps (PDRD_OPT_PIPELINE, the default of pdrd_iterate): the
nodes shared between patches are finished in the next sweep's gather and
by one fix-up after the last sweep. Every iteration count (the first two
sweeps eager, then pairs of sweeps from a HIP graph, then the rest eager)
gives bit for bit the state, residual, time step, wave speed and norm of a
fix-up launch after every sweep."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import CASES, quad_orders

pytestmark = pytest.mark.gpu


def _pair(name, scheme):
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    out = []
    for pipe in (1, 0):
        m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
        m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
        m.set_option(L.OPT_PIPELINE, pipe)
        assert m.get_option(L.OPT_PIPELINE) == pipe
        m.set_vec_function("solution", u)
        bn = mesh.boundaries.get("xmin", mesh.boundaries.get("Edge3"))
        if physics in (L.ADVECTION2D, L.BURGERS2D):
            m.add_boundary_condition("StrongDirichlet", "inflow", bn, u[bn])
        out.append(m)
    cfl = 0.3 if physics in (L.EULER2D, L.EULER3D) else 0.5
    return out, cfl


@pytest.mark.parametrize("name,scheme", [("burgers_p2", L.B), ("euler2d_mixed", L.N), ("euler3d_p1", L.LDA),
                                         ("adv_p3", L.PSI)])
def test_pipelined_sweeps_equal_fixup_per_sweep(name, scheme):
    (mp, mf), cfl = _pair(name, scheme)
    for n in (1, 2, 3, 4, 5, 8):
        print(f"pipeline {name} scheme {scheme}: {n} iterations", flush=True)
        for excl in (False, True):
            a = mp.iterate(cfl, n, exclude_dirichlet=excl)
            b = mf.iterate(cfl, n, exclude_dirichlet=excl)
            assert np.array_equal(a, b), (n, excl, a, b)
            for f in ("solution", "residuals", "time_step", "wave_speed"):
                assert np.array_equal(mp.vec_function(f), mf.vec_function(f)), (n, f)


def test_pipelined_timing_mode_matches():
    (mp, mf), cfl = _pair("burgers_p2", L.B)
    mp.set_option(L.OPT_TIMING, 1)
    for n in (1, 6):
        mp.iterate(cfl, n)
        mf.iterate(cfl, n)
        assert np.array_equal(mp.vec_function("solution"), mf.vec_function("solution"))
    c, nd = mp.last_timing()
    assert c > 0 and nd > 0
