"""GPU runs with remeshing (Forging3DSolver.Step's remesh trigger and recovery ladder, :540-638, RemeshCore
:1958-2102): the solver remeshes where the FP64 oracle does, onto the same mesh, and continues with the same loads;
a checkpoint of one mesh restores onto a new one (RestoreFrom :1899-1921)."""
import dataclasses

import numpy as np
import pytest

from impressionfea_amd.config import hexdie_production
from impressionfea_amd.forging3d import Forging3DSolver

pytestmark = pytest.mark.gpu


def _hexdie():
    wl = hexdie_production(nxy=8, nz=12)
    cfg = dataclasses.replace(wl.cfg, remesh=True)
    x, t, ox, oy = wl.mesh()
    return wl, cfg, x, t, ox, oy


def test_closed_die_remesh_matches_oracle(oracle_lib):
    """The small closed hex die (8 x 8 x 12 cells) with remeshing on: its initial Kuhn mesh (worst tet 0.093) is
    remeshed at the first increment, on the GPU solver as in the oracle, onto the same stuffed mesh bit for bit (the
    remesh runs before the solve, from identical states); the next increments' loads agree to 1e-6 (FP64) and neither
    side remeshes again."""
    import oracle as O
    wl, cfg, x, t, ox, oy = _hexdie()
    o = O.Oracle(cfg, x, t, ox, oy, wl.gap0)
    o.initial(float(x[:, 2].min()))
    o.set_remesh(cfg.remesh_quality)
    s = Forging3DSolver(cfg, x, t, ox, oy, wl.gap0)
    s.initial(float(x[:, 2].min()))
    for k in range(3):
        print(f"closed-die remesh run, increment {k + 1}", flush=True)
        ro = o.step()
        rs = s.step()
        assert rs.converged
        assert s.remesh_count == o.remesh_count() == 1
        assert (s.nn, s.ne) == (o.nn, o.ne)
        if k == 0:
            tt, a, b = o.mesh()
            assert np.array_equal(s.mesh.tets, tt) and np.array_equal(s.mesh.onX0, a)
        assert rs.load == pytest.approx(ro["load"], rel=1e-6)
        assert rs.gap == pytest.approx(ro["gap"], abs=1e-12)
    assert s.remesh_log[0][0] == 1 and s.remesh_log[0][1].startswith("element quality")
    assert s.remesh_log[0][2]["q_new"] > 0.2
    s.close()


def test_restore_onto_a_new_mesh(oracle_lib):
    """A checkpoint taken after two increments restores onto a solver built on another mesh: RestoreFrom remeshes the
    checkpoint's workpiece at this solver's h and remaps its history, and the run continues (the next load within 3 %
    of the uninterrupted run's; the meshes differ)."""
    wl, cfg, x, t, ox, oy = _hexdie()
    a = Forging3DSolver(cfg, x, t, ox, oy, wl.gap0)
    a.initial(float(x[:, 2].min()))
    a.step()
    a.step()
    cp = a.checkpoint()
    r_next = a.step()
    fixed = dataclasses.replace(cfg, remesh=False)
    wl2 = hexdie_production(nxy=6, nz=9)
    x2, t2, ox2, oy2 = wl2.mesh()
    b = Forging3DSolver(fixed, x2, t2, ox2, oy2, wl.gap0)
    b.restore_from(cp)
    assert b.remesh_count == 1 and b.step_index == cp.step
    assert b.gap == cp.gap
    rb = b.step()
    assert rb.converged
    assert rb.load == pytest.approx(r_next.load, rel=3e-2)
    a.close()
    b.close()
