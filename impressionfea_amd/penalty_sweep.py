"""Mixed precision across the penalty range on the small upset (the 10x10x10-cell Kuhn billet, 3,993 DOF).

For each penalty factor (1e4, 1e5, 1e6, 1e7 by default) the billet is stepped by the plain FP64 solver; at the start
of every increment the Newton system (Forging3DSolver.cs:965-1007) is also solved by the ISpdSolver backend in mixed
precision (FP32 MFMA supernode updates, FP64 diagonal POTRF, FP64 residual and iterative refinement to |r| <= 1e-10
|b|). Each row reports the refinement iterations, whether and where the FP64 fallback triggered (the FP32 factor not
positive definite, or refinement stalling), and the error against the FP64 factor's solution; a second solver steps
the same billet in mixed precision end to end and its ram load is compared with the FP64 one. The contact update
between increments is a discrete decision (a node is in contact or not): once the two steppers' contact sets differ,
their loads differ by that decision, not by the arithmetic, so rows record whether the discrete state still agrees.

    python -m impressionfea_amd.penalty_sweep [--increments 20] [--penalties 1e4,1e5,1e6,1e7]

prints one JSON object per increment and a summary per penalty factor.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import scipy.sparse as sp

from .config import upset_small
from .forging3d import Forging3DSolver
from .spd import HipSpdSolver


def sweep(penalties=(1e4, 1e5, 1e6, 1e7), increments=20, nxy=10, nz=10, refine_tol=1e-10, tolerance=None,
          log=print, keep_systems=False):
    """Every Newton system of both steppers' increments, solved by the slot in FP64 and in mixed precision; with
    keep_systems each row also holds its system (pattern, values, right-hand side) for an independent check."""
    out = []
    for pf in penalties:
        over = dict(penalty_factor=pf) if tolerance is None else dict(penalty_factor=pf, tolerance=tolerance)
        wl = upset_small(nxy=nxy, nz=nz, **over)
        ref = Forging3DSolver.from_workload(wl, precision=0)
        mix = Forging3DSolver.from_workload(wl, precision=1)
        s64 = HipSpdSolver(precision=0)
        s32 = HipSpdSolver(precision=1, refine_tol=refine_tol)
        s64.Analyze(ref.ndof, ref.colPtr, ref.rowIdx)
        s32.Analyze(ref.ndof, ref.colPtr, ref.rowIdx)
        rows = []
        same_state = True   # the two steppers have taken the same discrete decisions (contact set, step size) so far
        for k in range(increments):
            same_state &= bool(np.array_equal(ref.contact, mix.contact)) and ref.gap == mix.gap
            ref.prepare()
            val, b = ref.linear_system(newton=k > 0)
            A = sp.csc_matrix((val, ref.rowIdx, ref.colPtr), shape=(ref.ndof, ref.ndof))
            s64.Factorize(val)
            x64 = s64.Solve(b)
            fb0 = s32.stats()["fp64_fallbacks"]
            s32.Factorize(val)
            fb1 = s32.stats()["fp64_fallbacks"]
            x32 = s32.Solve(b)
            st = s32.stats()
            bn = np.linalg.norm(b)
            row = dict(penalty=pf, increment=k, refine_iterations=st["refine_iterations"],
                       fallback=("factor not SPD in FP32" if fb1 > fb0 else
                                 "refinement stalled" if st["fp64_fallbacks"] > fb1 else None),
                       rel_residual=float(np.linalg.norm(A @ x32 - b) / bn),
                       rel_residual_fp64=float(np.linalg.norm(A @ x64 - b) / bn),
                       err_vs_fp64=float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)))
            ra, rb = ref.step(), mix.step()
            row.update(load_fp64=ra.load, load_mixed=rb.load, load_rel_err=abs(rb.load / ra.load - 1),
                       step_fp64_fallbacks=rb.stats["fp64_fallbacks"], iterations_fp64=ra.iterations,
                       iterations_mixed=rb.iterations, converged_fp64=ra.converged, converged_mixed=rb.converged,
                       last_rel_fp64=ra.stats["last_rel"], last_rel_mixed=rb.stats["last_rel"],
                       pcg_iterations_fp64=ra.stats["pcg_iterations"], pcg_iterations_mixed=rb.stats["pcg_iterations"],
                       same_discrete_state=same_state,
                       contact_differences=int(np.count_nonzero(ref.contact != mix.contact)))
            log(json.dumps(row))
            if keep_systems:
                row["system"] = (ref.colPtr, ref.rowIdx, val, b)
            rows.append(row)
        for h in (s64, s32):
            h.Dispose()
        ref.close()
        mix.close()
        fb = [r["increment"] for r in rows if r["fallback"]]
        summ = dict(penalty=pf, increments=increments,
                    max_refine_iterations=max(r["refine_iterations"] for r in rows),
                    fallback_increments=fb, first_fallback=fb[0] if fb else None,
                    fallback_kinds=sorted({r["fallback"] for r in rows if r["fallback"]}),
                    max_err_vs_fp64=max(r["err_vs_fp64"] for r in rows),
                    max_rel_residual=max(r["rel_residual"] for r in rows),
                    max_load_rel_err=max(r["load_rel_err"] for r in rows),
                    max_load_rel_err_same_state=max((r["load_rel_err"] for r in rows if r["same_discrete_state"]),
                                                    default=None),
                    first_contact_difference=next((r["increment"] for r in rows if r["contact_differences"]), None),
                    step_fp64_fallbacks=sum(r["step_fp64_fallbacks"] for r in rows))
        log(json.dumps({"summary": summ}))
        out.append((summ, rows))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--increments", type=int, default=20)
    ap.add_argument("--penalties", default="1e4,1e5,1e6,1e7")
    ap.add_argument("--nxy", type=int, default=10)
    ap.add_argument("--nz", type=int, default=10)
    ap.add_argument("--tolerance", type=float, default=None, help="nonlinear tolerance (default: the workload's 1e-6)")
    a = ap.parse_args()
    sweep(tuple(float(p) for p in a.penalties.split(",")), a.increments, a.nxy, a.nz, tolerance=a.tolerance)


if __name__ == "__main__":
    main()
