"""Writes tests/golden/remesh_states.npz: the FP64 oracle's state of the small closed hex die (8 x 8 x 12 Kuhn cells,
h = 3.75 mm) with remeshing on, right before the increments in which it remeshes the second and third time (after
increments 26 and 49 of the run from the initial billet, flash forming). The parity tests remesh these states with the
product and the oracle and compare the new meshes and remapped fields bit for bit.

Run from the repository root: python tests/golden/make_remesh_states.py (about 6 minutes on 8 cores).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from impressionfea_amd.config import hexdie_production  # noqa: E402
from impressionfea_amd.placement import tet_volumes  # noqa: E402

SAVE_AFTER = (26, 49)


def main():
    O.build()
    wl = hexdie_production(nxy=8, nz=12)
    x, t, ox, oy = wl.mesh()
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.initial(float(x[:, 2].min()))
    o.set_remesh(0.12)
    vol0 = float(np.cumsum(tet_volumes(x, t))[-1]) * wl.cfg.sym_factor
    out = dict(vol0=np.array(vol0))
    for k in range(1, max(SAVE_AFTER) + 1):
        r = o.step()
        print(f"increment {k}: load {r['load']:.6e} gap {r['gap']:.4f} nn {o.nn} remeshes {o.remesh_count()}",
              flush=True)
        if k in SAVE_AFTER:
            st, es = o.state(), o.element_state()
            tt, a, b = o.mesh()
            pre = f"s{k}_"
            out.update({pre + "x": st["x"], pre + "tets": tt, pre + "on_x0": a, pre + "on_y0": b,
                        pre + "contact": st["contact"], pre + "vg": st["vg"], pre + "eps": st["eps"],
                        pre + "sig": st["elSig"], pre + "rate": st["elRate"], pre + "mean": es["elMean"],
                        pre + "dmg": es["dmg"], pre + "gap": np.array(st["gap"]),
                        pre + "remeshes": np.array(o.remesh_count())})
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "remesh_states.npz"), **out)


if __name__ == "__main__":
    main()
