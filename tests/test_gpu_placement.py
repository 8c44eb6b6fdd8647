"""GPU parity of TryUpdate's placement repair (Forging3DSolver.cs:1455-1537: contact placement, ImprovePlacement,
SmoothInterior, the collapse check) against the FP64 CPU oracle: the host-side repair, driven by the device's tet
quality kernel, gives the oracle's node positions bit for bit, the same contact sides and the same move counts."""
import numpy as np
import pytest

from impressionfea_amd.config import upset_small
from impressionfea_amd.forging3d import Forging3DSolver
from impressionfea_amd.mesh import kuhn_cylinder
from test_oracle_pins import placement_box_case

pytestmark = pytest.mark.gpu


def _compare(o, s, dt):
    why_o, moves = o.try_update(dt)
    why_s = s.try_update(dt)
    assert (why_o == "") == (why_s == "")
    assert tuple(s.last_repair) == tuple(moves)
    st = o.state()
    assert np.array_equal(s.x, st["x"])
    assert np.array_equal(s.contact, st["contact"])
    assert s.gap == st["gap"]
    return moves


def test_flat_die_box_repair_matches_oracle(oracle_lib):
    """The lagging top node and the rising node below it (one ImprovePlacement move, SmoothInterior moves)."""
    import oracle as O
    cfg, x, t, contact, vg, g0, dt, top, below = placement_box_case()
    z = np.zeros(len(x), bool)
    o = O.Oracle(cfg, x, t, z, z.copy(), g0)
    o.set_state(x=x, vg=vg, contact=contact, gap=g0)
    s = Forging3DSolver(cfg, x, t, z, z.copy(), g0)
    s.x, s.vg, s.contact, s.gap = x.copy(), vg.copy(), contact.copy(), g0
    print("flat-die box", flush=True)
    reverts, smooth = _compare(o, s, dt)
    assert reverts == 1 and smooth >= 1
    s.close()


@pytest.mark.parametrize("amp", [0.05, 0.06])
def test_noisy_cylinder_repair_matches_oracle(oracle_lib, amp):
    """The quarter upset cylinder (6 x 8 cells) from its initial state with the velocities perturbed by amp * h / dt
    (seeded): both repair passes act on several nodes of both dies and of the interior."""
    import oracle as O
    wl = upset_small(nxy=6, nz=8, quarter=True)
    x, t, ox, oy = kuhn_cylinder(wl.diameter, wl.height, wl.nxy, wl.nz, seed=3, quarter=wl.quarter)
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.initial(0.0)
    st = o.state()
    dt = 0.05
    vg = st["vg"] + amp * wl.cfg.h / dt * np.random.default_rng(5).standard_normal(st["vg"].shape)
    o.set_state(x=st["x"], vg=vg, contact=st["contact"], gap=st["gap"])
    s = Forging3DSolver(wl.cfg, x, t, ox, oy, wl.gap0)
    s.initial(0.0)
    assert np.array_equal(s.x, st["x"]) and np.array_equal(s.contact, st["contact"])
    s.vg = vg.copy()
    print(f"noisy cylinder amp={amp}", flush=True)
    reverts, smooth = _compare(o, s, dt)
    assert reverts >= 1 and smooth >= 1
    s.close()
