"""GPU parity of FinalEvaluation's element mean stress, AccumulateDamage and the converged velocities at the reference's
default penalty, against the FP64 CPU oracle, through the C ABI.

The element mean stress kPen * g_n . v is a product of a large penalty and a small divergence, so it is compared on
identical inputs (the same velocity field on both sides) to rounding, and along a stepped run at the reference's
default penalty factor 1e4 (Forging3DSolver.cs:27), where the FP64 solutions agree to ~1e-12.
"""
import ctypes as C

import numpy as np
import pytest

from impressionfea_amd import _native as N
from impressionfea_amd.config import upset_small
from impressionfea_amd.forging3d import Forging3DSolver
from impressionfea_amd.mesh import kuhn_cylinder

pytestmark = pytest.mark.gpu


def _pair(wl, seed=0):
    import oracle as O
    x, t, ox, oy = kuhn_cylinder(wl.diameter, wl.height, wl.nxy, wl.nz, seed=seed, quarter=wl.quarter)
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.initial(0.0)
    s = Forging3DSolver(wl.cfg, x, t, ox, oy, wl.gap0)
    s.initial(0.0)
    return o, s


def test_mean_stress_and_damage_same_state(oracle_lib):
    """One FinalEvaluation and one damage update on the same (perturbed, compressible) velocity field on both sides:
    the element mean stress to 1e-12 of kPen |g| |v|, the deviatoric stress to 1e-12 of its largest entry, and the
    damage to 1e-9 of its largest value (the trigonometric largest-eigenvalue formula loses digits near double roots,
    and the device's acos/cos differ from the host's in the last bits)."""
    wl = upset_small(nxy=6, nz=8, quarter=True)
    o, s = _pair(wl, seed=4)
    o.prepare()
    s.prepare()
    rng = np.random.default_rng(9)
    v = o.v() * (1 + 0.2 * rng.standard_normal(o.ndof)) + 0.02 * rng.standard_normal(o.ndof)
    o.set_v(v)
    o.final()
    s.set_v(v)
    da = N.i32(s.die_axis)
    pw, rc = C.c_double(0), C.c_double(0)
    assert N.lib().ifea_step_final(s._h, N.ptr(da, C.c_int), C.byref(pw), C.byref(rc), None, None, None) == 0
    eo = o.element_state()
    eg = s.element_results()
    kpen = o.constants()["kPen"]
    gscale = kpen * np.abs(o.geometry()["grad"]).max() * np.abs(v).max()
    assert np.abs(eg["elMean"] - eo["elMean"]).max() <= 1e-12 * gscale
    assert np.abs(eo["elMean"]).max() > 1e-6 * gscale          # a compressible field: the mean stress is not noise
    assert np.abs(eg["elDev"] - eo["elDev"]).max() <= 1e-12 * np.abs(eo["elDev"]).max()
    dt = 0.05
    o.accumulate_damage(dt)
    assert N.lib().ifea_step_advance(s._h, dt, None, None, None) == 0
    d_o, d_g = o.element_state()["dmg"], s.element_results()["dmg"]
    assert (d_o > 0).sum() > 0.1 * o.ne
    assert np.array_equal(d_o > 0, d_g > 0)
    assert np.abs(d_g - d_o).max() <= 1e-9 * d_o.max()
    # a restart value replaces the state; NULL clears it
    s.set_damage(np.full(o.ne, 0.25))
    assert np.all(s.element_results()["dmg"] == 0.25)
    s.set_damage(None)
    assert np.all(s.element_results()["dmg"] == 0.0)
    s.close()


def test_velocities_mean_stress_damage_default_penalty(oracle_lib):
    """Three increments of the small upset at the reference's default penalty factor 1e4 and a tight nonlinear
    tolerance: the converged velocities agree with the oracle's to 1e-9 relative (the plain-FP64 velocity bar), the
    ram load to 1e-6, and the element mean stress and accumulated damage of the stepped run to 1e-6 of their
    largest values. Past the Newton limit the reference iterates Picard steps, which converge linearly, so the
    iteration cap (SimulationConfig3D.MaxIterations) is raised for the 1e-12 tolerance on both sides."""
    wl = upset_small(nxy=8, nz=8, tolerance=1e-12, penalty_factor=1e4, max_iterations=300)
    o, s = _pair(wl)
    rows = []
    for k in range(3):
        ro, rs = o.step(), s.step()
        assert rs.converged
        vg = o.state()["vg"]
        dv = np.linalg.norm(s.vg - vg) / np.linalg.norm(vg)
        eo, eg = o.element_state(), s.element_results()
        dm = np.abs(eg["elMean"] - eo["elMean"]).max() / np.abs(eo["elMean"]).max()
        dd = np.abs(eg["dmg"] - eo["dmg"]).max() / max(eo["dmg"].max(), 1e-300)
        rows.append((k, rs.load / ro["load"] - 1, dv, dm, dd))
        print(f"increment {k}: load rel {rows[-1][1]:.2e} |dv|/|v| {dv:.2e} mean stress {dm:.2e} damage {dd:.2e}")
        assert rs.load == pytest.approx(ro["load"], rel=1e-6), rows
        assert dv <= 1e-9, rows
        assert dm <= 1e-6 and dd <= 1e-6, rows
        assert eo["dmg"].max() > 0
    s.close()


@pytest.mark.parametrize("workload", ["upset_quarter", "hexdie", "conrod"])
def test_device_frames_bit_equal(oracle_lib, workload):
    """PrepareStep's frames on the device (ifea_step_prepare_device) equal the oracle's PrepareStep (:669-727) and the
    host's compute_frames bit for bit: the quarter upset (symmetry planes with both flat dies), the hex die and the
    con-rod die (curved die surfaces). Every surface node is put in contact (upper half with the top die, lower half
    with the bottom die) and, for the hex die, the billet is widened to reach the cavity walls, so that the die normals
    of the curved parts are exercised; the die axes and fixed velocities are compared too."""
    import oracle as O
    from impressionfea_amd.config import conrod_production, hexdie_production
    from impressionfea_amd.forging3d import compute_frames
    from impressionfea_amd.mesh import TetMesh
    wl = {"upset_quarter": lambda: upset_small(quarter=True, nxy=5, nz=6),
          "hexdie": lambda: hexdie_production(nxy=12, nz=18),
          "conrod": lambda: conrod_production(nxy=8, nz=38)}[workload]()
    x, t, ox, oy = wl.mesh()
    if workload == "hexdie":
        x = x * np.array([1.3, 1.3, 1.0])
    surf = TetMesh(x.shape[0], t, ox, oy).is_surface_node
    zmid = 0.5 * (x[:, 2].min() + x[:, 2].max())
    contact = np.where(surf, np.where(x[:, 2] > zmid, 2, 1), 0).astype(np.int32)
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.set_state(x=x, contact=contact, gap=wl.gap0)
    o.prepare()
    fr = o.frames()
    s = Forging3DSolver(wl.cfg, x, t, ox, oy, wl.gap0)
    s.x, s.contact, s.gap = x.copy(), contact.copy(), wl.gap0
    s.prepare(np.zeros((s.nn, 3)))
    ax, nfix, fv = s.frames()
    assert np.array_equal(ax, fr["ax"])
    assert np.array_equal(nfix, fr["nFix"]) and np.array_equal(s.die_axis, fr["dieAxis"])
    assert np.array_equal(fv, fr["fixedVal"])
    hx = compute_frames(x, ox, oy, contact, wl.gap0, wl.cfg)
    assert np.array_equal(ax, hx[0]) and np.array_equal(fv, hx[3])
    if workload != "upset_quarter":
        # curved parts of the dies give tilted normals (the first local axis of a contact node is its die normal
        # when no symmetry plane comes first)
        assert np.abs(ax[contact > 0, 0, :2]).max() > 1e-2
    s.close()


def test_checkpoint_restart_continues_the_run(tmp_path):
    """Forging3DSolver.Checkpoint / RestoreFrom through the checkpoint file: a fresh solver restored from the state
    after two increments (with the trend history kept) steps the third increment like the uninterrupted run. The two
    take different Newton paths (the restarted solver has no cached factor), and the velocities converge to the
    configured 1e-6 relative change, so the bars are: ram load 1e-6, positions 1e-6 of their scale, strain (which
    integrates the strain rate) 1e-5 of its largest value, and damage 1e-3: it also integrates the mean stress
    kPen g . v, in which the penalty magnifies the velocity's convergence tolerance. The run's telemetry rows go to a
    CSV in the reference's format and read back."""
    from impressionfea_amd import formats as F
    wl = upset_small(nxy=6, nz=8)
    a = Forging3DSolver.from_workload(wl)
    v0 = float(F.tet_volumes(a.x, a.mesh.tets).sum())
    rows, cp_path = [], str(tmp_path / "after2.npz")
    for k in range(3):
        r = a.step()
        rows.append(F.telemetry_row(a, r, samples=F.cavity_samples(wl.cfg), v0=v0))
        if k == 1:
            F.save_checkpoint(cp_path, a.checkpoint())
    ref = dict(load=r.load, x=a.x.copy(), eps=a.eps.copy(), dmg=a.element_results()["dmg"])
    b = Forging3DSolver.from_workload(wl)
    b.restore_from(F.load_checkpoint(cp_path), warm=True)
    rb = b.step()
    print(f"restart: load {rb.load:.9e} vs {ref['load']:.9e} N; Newton {rb.iterations} vs {r.iterations}")
    assert rb.step == 3 and rb.load == pytest.approx(ref["load"], rel=1e-6)
    assert np.abs(b.x - ref["x"]).max() <= 1e-6 * np.abs(ref["x"]).max()
    assert np.abs(b.eps - ref["eps"]).max() <= 1e-5 * ref["eps"].max()
    d = b.element_results()["dmg"]
    assert ref["dmg"].max() > 0 and np.abs(d - ref["dmg"]).max() <= 1e-3 * ref["dmg"].max()
    csv = str(tmp_path / "run.csv")
    F.write_telemetry_csv(csv, rows, comments=[wl.name])
    back, _ = F.read_telemetry_csv(csv)
    assert [q.step for q in back] == [1, 2, 3]
    assert back[-1].load_tonnes == pytest.approx(r.load / F.TONNE_N, abs=1e-3)
    assert back[-1].stroke == pytest.approx(wl.gap0 - a.gap, abs=1e-4) and abs(back[-1].volume_change_percent) < 1.0
    a.close()
    b.close()
