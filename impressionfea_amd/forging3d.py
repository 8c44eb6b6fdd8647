"""Host driver of the MI355X solve step: the increment loop of the reference's Forging3DSolver on a fixed mesh.

Mirrors src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:
  constructor constants and initial contact/velocity   :125-187
  Step (one stroke increment and its recovery ladder)   :540-638 (remesh: remesh.py)
  PrepareStep host part: nodal frames and fixed DOFs    :660-727   (vectorised numpy; die geometry is host side)
  LayerLimitedIncrement                                 :236-249
  TryUpdate (contact, placement and its repair)         :1455-1537, ImprovePlacement :1625-1679, SmoothInterior
                                                        :1776-1838 (tet quality on the device)
Everything between PrepareStep's frames and the strain update runs on the GPU through include/ifea_step.h.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _native as N
from . import placement as PL
from . import remesh as RM
from .placement import tet_quality
from .config import StepConfig, Workload
from .errors import MeshDistortionError, SolverDivergenceError, raise_for
from .mesh import TetMesh, build_pattern

FREE, BOTTOM, TOP = 0, 1, 2


def _dot3(a, b):
    """Row-wise a . b as (a0 b0 + a1 b1) + a2 b2: the reference's Vec3.Dot order, never a fused or reassociated sum."""
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _normalized(a):
    return a / np.sqrt(_dot3(a, a))[..., None]


def compute_frames(x, onX0, onY0, contact, gap, cfg: StepConfig):
    """PrepareStep's nodal frames (Forging3DSolver.cs:669-727), vectorised over nodes; the arithmetic of every node is
    the reference's (and the device kernel's) in the same order.

    Returns ax [nn,3,3] (rows ax0, ax1, ax2), nfix [nn], dieAxis [nn], fixedVal [3nn].
    """
    nn = x.shape[0]
    dirs = np.zeros((nn, 3, 3))
    vals = np.zeros((nn, 3))
    k = np.zeros(nn, dtype=np.int64)
    die_idx = np.full(nn, -1, dtype=np.int64)
    ar = np.arange(nn)
    m = onX0.astype(bool)
    dirs[ar[m], k[m], 0] = 1.0
    k[m] += 1
    m = onY0.astype(bool)
    dirs[ar[m], k[m], 1] = 1.0
    k[m] += 1
    for side, die, off in ((BOTTOM, cfg.bottom, 0.0), (TOP, cfg.top, gap)):
        m = contact == side
        if not m.any():
            continue
        _, nrm = die.project(x[m], off)
        idx = ar[m]
        dirs[idx, k[m]] = nrm
        dvel = np.array([0.0, 0.0, -cfg.ram]) if side == TOP else np.zeros(3)
        vals[idx, k[m]] = _dot3(nrm, dvel)
        die_idx[idx] = k[m]
        k[m] += 1
    # nodes without constraints keep the identity frame: the Gram-Schmidt below runs on the constrained ones only
    full_e, full_ev = np.broadcast_to(np.eye(3), (nn, 3, 3)).copy(), np.zeros((nn, 3))
    full_mc, full_da = np.zeros(nn, dtype=np.int64), np.full(nn, -1, dtype=np.int64)
    sub = np.nonzero(k > 0)[0]
    dirs, vals, k, die_idx = dirs[sub], vals[sub], k[sub], die_idx[sub]
    nn_all, nn = nn, sub.size
    ar = np.arange(nn)
    e = np.zeros((nn, 3, 3))
    ev = np.zeros((nn, 3))
    mc = np.zeros(nn, dtype=np.int64)
    die_axis = np.full(nn, -1, dtype=np.int64)
    for q in range(3):
        active = q < k
        d = dirs[:, q].copy()
        v = vals[:, q].copy()
        for s in range(3):
            sm = active & (s < mc)
            proj = _dot3(d, e[:, s])
            d[sm] -= proj[sm, None] * e[sm, s]
            v[sm] -= proj[sm] * ev[sm, s]
        ln = np.sqrt(_dot3(d, d))
        keep = active & (ln >= 0.2)
        ki = ar[keep]
        e[ki, mc[keep]] = (1.0 / ln[keep])[:, None] * d[keep]
        ev[ki, mc[keep]] = v[keep] / ln[keep]
        hit = keep & (die_idx == q)
        die_axis[hit] = mc[hit]
        mc[keep] += 1
    # completion to an orthonormal frame
    m0 = mc == 0
    e[m0] = np.eye(3)
    m1 = mc == 1
    if m1.any():
        e0 = e[m1, 0]
        helper = np.where((np.abs(e0[:, 0]) < 0.8)[:, None], np.array([1.0, 0, 0]), np.array([0, 1.0, 0]))
        e1 = np.cross(e0, helper)
        e1 = _normalized(e1)
        e[m1, 1] = e1
        e[m1, 2] = np.cross(e0, e1)
    m2 = mc == 2
    if m2.any():
        e2 = np.cross(e[m2, 0], e[m2, 1])
        e[m2, 2] = _normalized(e2)
    full_e[sub], full_ev[sub], full_mc[sub], full_da[sub] = e, ev, mc, die_axis
    e, ev, mc, die_axis, nn = full_e, full_ev, full_mc, full_da, nn_all
    fixed_val = np.zeros(3 * nn)
    fv = fixed_val.reshape(nn, 3)
    for q in range(3):
        mq = q < mc
        fv[mq, q] = ev[mq, q]
    return e, mc.astype(np.int32), die_axis.astype(np.int32), fixed_val


@dataclass
class IncrementResult:
    step: int
    load: float
    reaction: float
    iterations: int
    converged: bool
    dt: float
    gap: float
    stats: dict
    ms_solve: float      # PrepareStep + SolveVelocity + FinalEvaluation + strain update (the unit of work)
    ms_update: float     # TryUpdate (host geometry), timed separately


class Forging3DSolver:
    """One billet stepped by increments on the GPU (Forging3DSolver.Step); remeshing when cfg.remesh is set."""

    def __init__(self, cfg: StepConfig, x, tets, onX0, onY0, gap0: float, precision: int = 0, perm=None,
                 warm_start=None, first_solve_policy: bool = False, device_frames: bool = True,
                 pcg_tol: float = 1e-7):
        self.cfg = cfg
        consts = cfg.constants()
        self.rate_ref, self.rate_cut, self.sig_ref = consts["rateRef"], consts["rateCut"], consts["sigRef"]
        self.k_pen, self.u0 = consts["kPen"], consts["u0"]
        self.gap = self.gap0 = float(gap0)
        self.ds = cfg.stroke_increment
        self.warm = False
        self.step_index = 0
        self.n_converged = 0     # converged velocity fields so far (FinalEvaluation calls)
        self.time = 0.0
        self.warm_start = warm_start
        self.vg_prev = None
        self.recovery_pass = 0     # Step's recovery ladder (:553-621): 0 as is, 1 after a remesh
        self.precision = precision
        self._first_solve = bool(first_solve_policy)
        # relative residual at which PCG on the cached factor stops: the reference's 1e-7 (SolveLinear,
        # Forging3DSolver.cs:984) unless an inexact Newton iteration is asked for (a looser tolerance; Newton's own
        # convergence test on the velocity step is unchanged)
        self.pcg_tol = float(pcg_tol)
        # PrepareStep's frames on the device (the host's compute_frames is the same arithmetic, kept for checks)
        self.device_frames = bool(device_frames)
        self._h = None
        self._setup_mesh(x, tets, onX0, onY0, perm)
        # the initial volume (:184) and the model's symmetry planes, kept across remeshes
        self.vol0 = float(np.cumsum(PL.tet_volumes(self.x, self.mesh.tets))[-1]) * cfg.sym_factor
        self.symX, self.symY = bool(np.any(self.mesh.onX0)), bool(np.any(self.mesh.onY0))
        self.remesh_count = 0
        self.remesh_log = []       # (increment, reason, stats or None when the new mesh was rejected)
        self.ms_remesh = 0.0

    def _setup_mesh(self, x, tets, onX0, onY0, perm=None):
        """Allocate (:403-438): the mesh, its pattern and local lengths, and a device step for them; per-node and
        per-element state starts empty."""
        cfg = self.cfg
        self.close()
        self.x = np.array(x, dtype=np.float64)
        self.mesh = TetMesh(self.x.shape[0], tets, np.asarray(onX0, bool), np.asarray(onY0, bool))
        m = self.mesh
        self.nn, self.ne, self.ndof = m.nn, m.ne, 3 * m.nn
        self.nbr_start, self.nbr, self.colPtr, self.rowIdx = build_pattern(m)
        self.nnz = int(self.colPtr[-1])
        self.h_local = PL.local_lengths(m, self.x, cfg.h)   # LocalLengths (:443-466)
        L = N.lib()
        faces = m.boundary_faces[~m.face_is_symmetry]
        owners = m.boundary_face_owner[~m.face_is_symmetry]
        self._keep = [N.i32(m.tets), N.i32(self.colPtr), N.i32(self.rowIdx), N.i32(faces), N.i32(owners)]
        h = C.c_void_p()
        rc = L.ifea_step_create(C.byref(h), self.nn, self.ne, N.ptr(self._keep[0], C.c_int), N.ptr(self._keep[1], C.c_int),
                                N.ptr(self._keep[2], C.c_int), int(faces.shape[0]), N.ptr(self._keep[3], C.c_int),
                                N.ptr(self._keep[4], C.c_int))
        self._check(rc, "ifea_step_create")
        self._h = h
        self.set_params()
        if perm is not None:
            pm = N.i32(perm)
            self._check(L.ifea_step_set_perm(self._h, N.ptr(pm, C.c_int)), "set_perm")
        if self._first_solve:
            self._check(L.ifea_step_set_option(self._h, N.IFEA_STEP_OPT_FIRST_SOLVE, 1.0), "set_option")
        self._sym_flags = [np.ascontiguousarray(m.onX0, np.uint8), np.ascontiguousarray(m.onY0, np.uint8)]
        bp, tp = N.f64(np.asarray(cfg.bottom.p, float)), N.f64(np.asarray(cfg.top.p, float))
        self._check(L.ifea_step_set_geometry(self._h, N.ptr(self._sym_flags[0], C.c_ubyte),
                                             N.ptr(self._sym_flags[1], C.c_ubyte),
                                             int(cfg.bottom.kind), N.ptr(bp, C.c_double), int(cfg.top.kind),
                                             N.ptr(tp, C.c_double)), "set_geometry")
        self._host_frames = None
        # state
        self.contact = np.zeros(self.nn, dtype=np.int32)
        self.eps = np.zeros(self.ne)
        self.el_sig = np.full(self.ne, self.sig_ref)
        self.el_rate = np.zeros(self.ne)
        self.vg = np.zeros((self.nn, 3))
        self.release = np.zeros(self.nn, dtype=bool)
        self.last_repair = (0, 0)          # nodes moved by ImprovePlacement and SmoothInterior in the last TryUpdate
        self._state_dirty = True

    @staticmethod
    def _check(rc, what):
        if rc:
            raise_for(rc, what, N.last_error())

    def set_params(self, precision: int | None = None, profile: bool | None = None):
        if precision is not None:
            self.precision = precision
        if profile is not None:
            self.profile = bool(profile)
        c = self.cfg
        p = N.StepParams()
        p.C, p.n, p.m, p.eps0 = c.C, c.n, c.m, c.eps0
        p.friction, p.ram, p.kpen, p.rate_cut, p.u0 = c.friction, c.ram, self.k_pen, self.rate_cut, self.u0
        p.newton_blend, p.tolerance, p.pcg_tol = c.newton_blend, c.tolerance, self.pcg_tol
        p.max_iterations, p.pcg_budget = c.max_iterations, c.pcg_budget
        p.quadratic_ls, p.adaptive_newton = int(c.quadratic_ls), int(c.adaptive_newton)
        p.precision, p.newton_limit = int(self.precision), 20
        p.profile = int(getattr(self, "profile", False))
        self._check(N.lib().ifea_step_set_params(self._h, C.byref(p)), "set_params")

    @classmethod
    def from_workload(cls, wl: Workload, precision: int = 0, **kw):
        x, tets, ox, oy = wl.mesh()
        s = cls(wl.cfg, x, tets, ox, oy, wl.gap0, precision=precision, **kw)
        s.initial(zmin=float(x[:, 2].min()))
        return s

    # ------------------------------------------------------------------ constructor tail (:158-182)
    def initial(self, zmin: float, cx: float = 0.0, cy: float = 0.0):
        cfg, m = self.cfg, self.mesh
        tol = 0.05 * cfg.h
        surf = m.is_surface_node
        cb = cfg.bottom.clearance(self.x, 0.0)
        ct = cfg.top.clearance(self.x, self.gap)
        self.contact[:] = FREE
        bot = surf & (cb < tol)
        self.contact[bot] = BOTTOM
        topm = surf & ~bot & (ct < tol)
        self.contact[topm] = TOP
        for side, clr in ((BOTTOM, cb), (TOP, ct)):
            if not (self.contact == side).any():
                cand = np.where(surf, clr, np.inf)
                self.contact[int(np.argmin(cand))] = side
        self._place_contact_nodes()
        r = self.rate_ref
        p = self.x
        self.vg = np.stack([0.5 * r * (p[:, 0] - cx), 0.5 * r * (p[:, 1] - cy), -r * (p[:, 2] - zmin)], axis=1)

    def _sym(self, p):
        p = p.copy()
        p[self.mesh.onX0, 0] = 0.0
        p[self.mesh.onY0, 1] = 0.0
        return p

    def _place_contact_nodes(self):
        for side, die, off in ((BOTTOM, self.cfg.bottom, 0.0), (TOP, self.cfg.top, self.gap)):
            m = self.contact == side
            if m.any():
                q, _ = die.project(self.x[m], off)
                self.x[m] = q
        self.x = self._sym(self.x)

    # ------------------------------------------------------------------ increment (:540-638)
    def layer_limited_increment(self):
        m = self.contact == TOP
        if not m.any():
            return np.inf
        xs = self.x[m]
        zb = self.cfg.bottom.height(xs[:, 0], xs[:, 1])
        zt = self.cfg.top.height(xs[:, 0], xs[:, 1]) + self.gap
        ok = np.isfinite(zb) & np.isfinite(zt) & (zt > zb)
        if not ok.any():
            return np.inf
        return max(0.1 * float(np.min(zt[ok] - zb[ok])), self.cfg.stroke_increment / 32)

    def prepare(self, vguess=None):
        """PrepareStep: the nodal frames (on the device, or by compute_frames on the host), then the device part
        (geometry, ANP, friction, warm start)."""
        if vguess is None:
            trend = self.cfg.extrapolate_start and self.warm and self.vg_prev is not None
            vguess = 2 * self.vg - self.vg_prev if trend else self.vg
        L = N.lib()
        if self._state_dirty:
            e, s = N.f64(self.eps), N.f64(self.el_sig)
            self._check(L.ifea_step_set_element_state(self._h, N.ptr(e, C.c_double), N.ptr(s, C.c_double)), "state")
            self._state_dirty = False
        xs, ct, vgs = N.f64(self.x), N.i32(self.contact), N.f64(vguess)
        if self.device_frames:
            rc = L.ifea_step_prepare_device(self._h, N.ptr(xs, C.c_double), N.ptr(ct, C.c_int), N.ptr(vgs, C.c_double))
            self._check(rc, "PrepareStep")
            self._host_frames = None
            da = np.zeros(self.nn, np.int32)
            self._check(L.ifea_step_get_frames(self._h, None, None, N.ptr(da, C.c_int), None), "frames")
            self.die_axis = da
            return
        ax, nfix, die_axis, fixed_val = compute_frames(self.x, self.mesh.onX0, self.mesh.onY0, self.contact, self.gap,
                                                       self.cfg)
        self._host_frames = (ax, nfix, fixed_val)
        self.die_axis = die_axis
        axs, nf, fv = N.f64(ax), N.i32(nfix), N.f64(fixed_val)
        rc = L.ifea_step_prepare(self._h, N.ptr(xs, C.c_double), N.ptr(axs, C.c_double), N.ptr(nf, C.c_int),
                                 N.ptr(fv, C.c_double), N.ptr(ct, C.c_int), N.ptr(vgs, C.c_double))
        self._check(rc, "PrepareStep")

    def frames(self):
        """The current increment's frames (ax [nn,3,3], nfix [nn], fixed values [3nn]), read from the device after a
        device PrepareStep."""
        if self._host_frames is None:
            ax, nf, fv = np.zeros((self.nn, 3, 3)), np.zeros(self.nn, np.int32), np.zeros(self.ndof)
            self._check(N.lib().ifea_step_get_frames(self._h, N.ptr(ax, C.c_double), N.ptr(nf, C.c_int), None,
                                                     N.ptr(fv, C.c_double)), "frames")
            self._host_frames = (ax, nf, fv)
        return self._host_frames

    @property
    def ax(self):
        return self.frames()[0]

    @property
    def nfix(self):
        return self.frames()[1]

    @property
    def fixed_val(self):
        return self.frames()[2]

    def load_guess(self, vg):
        """LoadVelocityGuess (:854-869) of a global nodal velocity field with the current frames, on the device."""
        v = N.f64(vg)
        self._check(N.lib().ifea_step_load_guess(self._h, N.ptr(v, C.c_double)), "load_guess")

    def _learned_start(self) -> str:
        """PrepareStep with the learned guess, judged against the reference's own starts (the previous increment's
        field and its trend 2 v_n - v_{n-1}, ExtrapolateStart) by the convex functional Pi, the solver's own energy:
        the lowest-energy start is kept, so a poor prediction costs residual-only assemblies and never a worse start
        than the best of the reference's."""
        from .mlp import LearnedWarmStart, trend_start
        trend = trend_start(self)
        # the network's guess is the trend start plus its correction: the trend computed once for both
        guess = self.warm_start(self, trend) if isinstance(self.warm_start, LearnedWarmStart) else self.warm_start(self)
        self.prepare(guess)
        cands = {"learned": guess, "previous": self.vg, "trend": trend}
        best, best_e = None, np.inf
        for name, vg in cands.items():
            self.load_guess(vg)
            e = self.energy()
            if e < best_e:
                best, best_e = name, e
        if best != name:              # the device holds the last candidate evaluated
            self.load_guess(cands[best])
        return best

    def first_collapsed(self, p, qmin: float, quality=None) -> int:
        """Index of the first tet whose TetMesh.Quality at the node positions p is <= qmin, or -1 (the collapse check
        of TryUpdate, Forging3DSolver.cs:1509-1516), evaluated on the device; per-tet values into `quality` if given."""
        ps = N.f64(p)
        first = C.c_int(-1)
        qp = N.ptr(quality, C.c_double) if quality is not None else None
        self._check(N.lib().ifea_step_quality(self._h, N.ptr(ps, C.c_double), float(qmin), C.byref(first), qp),
                    "quality")
        return first.value

    def solve_velocity(self):
        st = N.StepStats()
        self._check(N.lib().ifea_step_solve_velocity(self._h, int(self.warm), C.byref(st)), "SolveVelocity")
        return st.as_dict()

    def final_evaluation(self):
        pw, rc = C.c_double(0), C.c_double(0)
        vg = np.zeros((self.nn, 3))
        rel = np.zeros(self.nn, np.uint8)
        da = N.i32(self.die_axis)
        self._check(N.lib().ifea_step_final(self._h, N.ptr(da, C.c_int), C.byref(pw), C.byref(rc),
                                            N.ptr(vg, C.c_double), None, N.ptr(rel, C.c_ubyte)), "FinalEvaluation")
        self.vg_prev = self.vg
        self.vg = vg
        self.n_converged += 1
        self.release = rel.astype(bool)
        return pw.value, rc.value

    # ------------------------------------------------------------------ geometry update and placement repair
    def _bad_tets(self, x, q_min):
        """Tets with TetMesh.Quality < q_min at x: the device evaluates every tet, the host re-evaluates those near the
        threshold with the reference's arithmetic so the decision is the reference's."""
        q = np.zeros(self.ne)
        self.first_collapsed(x, q_min, q)
        cand = np.nonzero(q < q_min + 1e-9)[0]
        if cand.size == 0:
            return cand
        return cand[tet_quality(x, self.mesh.tets[cand]) < q_min]

    def surface_normals(self, x):
        return PL.surface_normals(self.mesh, x)

    def improve_placement(self, x, contact, pred, nrm, gap, q_min=0.05):
        return PL.improve_placement(self.mesh, x, contact, pred, nrm, gap, q_min, self.h_local, self.cfg,
                                    self._bad_tets, self.recovery_pass)

    def smooth_interior(self, x, q_target=0.05, sweeps=6):
        return PL.smooth_interior(self.mesh, x, q_target, sweeps, self._bad_tets)

    def try_update(self, dt: float) -> str:
        """TryUpdate (:1455-1537): predicted positions, contact and placement, the placement repair (ImprovePlacement,
        SmoothInterior) and the collapse check on the device; '' on success, else the reason."""
        cfg = self.cfg
        gapT = max(0.0, self.gap - cfg.ram * dt)
        if gapT < 1e-9:
            gapT = 0.0
        p = self._sym(self.x + dt * self.vg)
        pred = p.copy()
        cT = np.zeros(self.nn, np.int32)
        # nodes in contact (and not released) stay on their die unless they left it
        for side, die in ((BOTTOM, cfg.bottom), (TOP, cfg.top)):
            off = gapT if side == TOP else 0.0
            m = (self.contact == side) & ~self.release
            if m.any():
                clr = die.clearance(p[m], off)
                stay = clr <= 0.05 * self.h_local[m]
                idx = np.where(m)[0]
                q, _ = die.project(p[idx[stay]], off)
                p[idx[stay]] = q
                cT[idx[stay]] = side
        free = ~(((self.contact != FREE) & ~self.release)) & self.mesh.is_surface_node
        cb = cfg.bottom.clearance(p, 0.0)
        ctp = cfg.top.clearance(p, gapT)
        p_clr = p.copy()
        hit = free & ((cb < 0) | (ctp < 0))
        max_pen = 0.0
        if hit.any():
            idx = np.where(hit)[0]
            to_bottom = cb[idx] <= ctp[idx]
            for side, die, off, sel in ((BOTTOM, cfg.bottom, 0.0, to_bottom), (TOP, cfg.top, gapT, ~to_bottom)):
                ii = idx[sel]
                if ii.size == 0:
                    continue
                q, _ = die.project(p[ii], off)
                pen = np.linalg.norm(q - p[ii], axis=1) / self.h_local[ii]
                max_pen = max(max_pen, float(pen.max()))
                p[ii] = q
                cT[ii] = side
        p = self._sym(p)
        if max_pen > 0.5:
            return "die penetration exceeds half the local element size"
        # both repair passes act only on tets below 0.05 (and the collapse check on tets below 0.005): one device
        # evaluation, with a margin for the last bits, tells whether there is any
        if self.first_collapsed(p, 0.05 + 1e-9) >= 0:
            nrm = self.surface_normals(self.x)
            self.last_repair = (self.improve_placement(p, cT, pred, nrm, gapT, 0.05), self.smooth_interior(p, 0.05, 6))
            bad = self.first_collapsed(p, 0.005)
        else:
            self.last_repair = (0, 0)
            bad = -1
        if bad >= 0:
            return f"tetrahedron {bad} collapsed"
        self.x = p
        self.contact = cT
        self.gap = gapT
        # both die clearances of the new positions (the learned start's features read them): those computed above,
        # recomputed where a node moved since
        moved = np.any(p != p_clr, axis=1)
        if moved.any():
            cb[moved] = cfg.bottom.clearance(p[moved], 0.0)
            ctp[moved] = cfg.top.clearance(p[moved], gapT)
        self.clearance = (self.x, cb, ctp)
        self.time += dt
        return ""

    # ------------------------------------------------------------------ remeshing (:1867-1880, :1945-2102)
    def remesh_trigger(self):
        """RemeshTrigger (:1867-1880): interior smoothing of the current mesh, then the reason to remesh, or None."""
        return RM.remesh_trigger(self.mesh, self.x, self.cfg.remesh_quality, self.cfg.h)

    def remesh(self, reason: str) -> bool:
        """Remesh / RemeshCore (:1945-2102): a stuffed mesh of the current workpiece with contact, placement repair and
        the state remapped; a new device step for it (its pattern analysed at the next factorisation). False when the
        new mesh is rejected (the state stays on the old one). The remapped element mean stress is not put on the
        device: the next FinalEvaluation writes it before it is read."""
        t0 = time.perf_counter()
        er = self.element_results()
        state = dict(eps=self.eps, dmg=er["dmg"], sig=self.el_sig, rate=self.el_rate, mean=er["elMean"], vg=self.vg)
        res = RM.remesh_core(self.cfg, self.mesh, self.x, state, self.gap, self.vol0, self.cfg.sym_factor,
                             self.sig_ref, self.symX, self.symY, self.recovery_pass)
        if res is None:
            self.remesh_log.append((self.step_index, reason, None))
            self.ms_remesh += 1e3 * (time.perf_counter() - t0)
            return False
        self._setup_mesh(res.x, res.tets, res.onX0, res.onY0)
        self.contact = np.asarray(res.contact, np.int32).copy()
        self.vg = res.vg
        self.set_element_state(res.eps, res.sig)
        self.el_rate = np.asarray(res.rate, np.float64)
        self.set_damage(res.dmg)
        self.warm = False
        self.vg_prev = None
        self.remesh_count += 1
        self.remesh_log.append((self.step_index, reason, res.stats))
        self.ms_remesh += 1e3 * (time.perf_counter() - t0)
        return True

    def step(self) -> IncrementResult:
        """One stroke increment (Forging3DSolver.Step :540-638). With cfg.remesh the remesh trigger runs first and a
        failed geometry update (or a stalled velocity solution) is retried once on a new mesh with the corner allowance
        of the first recovery pass (the ladder's refinement passes are not built); otherwise the mesh stays fixed."""
        cfg = self.cfg
        self.step_index += 1
        self.ms_remesh = 0.0
        if cfg.remesh:
            trig = self.remesh_trigger()
            if trig:
                self.remesh(trig)
        passes = 2 if cfg.remesh else 1
        ok, why = False, ""
        t0 = time.perf_counter()
        ms_pass = 0.0
        for pas in range(passes):
            self.recovery_pass = pas
            t0 = time.perf_counter()
            judged = None
            if self.warm_start is not None and self.warm:
                judged = self._learned_start()
            else:
                self.prepare()
            st = self.solve_velocity()
            trend = cfg.extrapolate_start and self.warm and self.vg_prev is not None
            st["start"] = judged or ("trend" if trend else "previous")
            if not st["converged"] and st["last_rel"] >= 1e-2:
                if pas < passes - 1 and self.remesh("stalled velocity solution"):
                    continue
                self.recovery_pass = 0
                raise SolverDivergenceError(f"Velocity solution did not converge at step {self.step_index}: "
                                            f"|dv|/|v| = {st['last_rel']:.2E}.")
            power, reaction = self.final_evaluation()
            load = power / cfg.ram * cfg.sym_factor
            reaction *= cfg.sym_factor
            t1 = time.perf_counter()
            ds = min(min(self.ds, cfg.stroke_increment), self.layer_limited_increment())
            dt = min(ds, self.gap) / cfg.ram
            if self.gap - cfg.ram * dt < 0.05 * ds:
                dt = self.gap / cfg.ram
            for attempt in range(6):
                dta = dt / (1 << attempt)
                why = self.try_update(dta)
                if not why:
                    dt = dta
                    ok = True
                    break
            if ok or pas == passes - 1:
                break
            ms_pass += 1e3 * (t1 - t0)
            self.remesh("fold during update")   # a rejected remesh still leaves the next pass
        self.recovery_pass = 0
        if not ok:
            raise MeshDistortionError(f"Mesh distortion at step {self.step_index}: {why}.")
        t2 = time.perf_counter()
        eps = np.zeros(self.ne)
        sig = np.zeros(self.ne)
        rate = np.zeros(self.ne)
        self._check(N.lib().ifea_step_advance(self._h, dt, N.ptr(eps, C.c_double), N.ptr(sig, C.c_double),
                                              N.ptr(rate, C.c_double)), "advance")
        self.eps, self.el_sig, self.el_rate = eps, sig, rate
        self.warm = True
        t3 = time.perf_counter()
        st["repair"] = tuple(self.last_repair)
        st["remesh_ms"] = self.ms_remesh
        st["remeshes"] = self.remesh_count
        return IncrementResult(self.step_index, load, reaction, st["iterations"], bool(st["converged"]), dt, self.gap,
                               st, ms_pass + 1e3 * ((t1 - t0) + (t3 - t2)), 1e3 * (t2 - t1))

    # ------------------------------------------------------------------ parity helpers
    def assemble(self, v, matrix=True, newton=True):
        e, pw = C.c_double(0), C.c_double(0)
        r = np.zeros(self.ndof)
        val = np.zeros(self.nnz) if matrix else None
        vv = N.f64(v) if v is not None else None
        self._check(N.lib().ifea_step_assemble(self._h, int(matrix), int(newton), N.ptr(vv, C.c_double), C.byref(e),
                                               C.byref(pw), N.ptr(r, C.c_double), N.ptr(val, C.c_double)), "Assemble")
        # the assembly overwrote the device's element flow stresses; the next PrepareStep restores the increment's
        # own (host copies) before it computes the friction coefficients from them
        self._state_dirty = True
        return e.value, pw.value, r, val

    def energy(self) -> float:
        """The functional Pi at the current device velocity (a residual-only Assemble, nothing copied back)."""
        e, pw = C.c_double(0), C.c_double(0)
        self._check(N.lib().ifea_step_assemble(self._h, 0, 0, None, C.byref(e), C.byref(pw), None, None), "Assemble")
        self._state_dirty = True
        return e.value

    def linear_system(self, newton: bool = True):
        """The Newton/Picard system of the current velocity as SolveVelocity forms it (Forging3DSolver.cs:965-1007):
        tangent and residual from Assemble, fixed local DOFs replaced by identity rows/columns, 1e-10 * max diagonal
        added to the free diagonal; returns (values in the CSC pattern, right-hand side -r)."""
        _, _, r, val = self.assemble(None, matrix=True, newton=newton)
        fixed = (np.arange(3)[None, :] < self.nfix[:, None]).ravel()
        col = np.repeat(np.arange(self.ndof), np.diff(self.colPtr))
        row = self.rowIdx
        msk = fixed[row] | fixed[col]
        val[msk] = np.where(row[msk] == col[msk], 1.0, 0.0)
        dg = (row == col) & ~fixed[col]
        val[dg] += 1e-10 * val[dg].max()
        return val, -np.where(fixed, 0.0, r)

    def element_results(self):
        """Element mean stress of the last FinalEvaluation (:1440-1450), accumulated Cockcroft-Latham damage
        (AccumulateDamage :1844-1855) and the deviatoric stress [ne, 6] of the last assembly, from the device."""
        mean, dmg, dev = np.zeros(self.ne), np.zeros(self.ne), np.zeros((self.ne, 6))
        self._check(N.lib().ifea_step_element_results(self._h, N.ptr(mean, C.c_double), N.ptr(dmg, C.c_double),
                                                      N.ptr(dev, C.c_double)), "element_results")
        return dict(elMean=mean, dmg=dmg, elDev=dev)

    def set_damage(self, dmg=None):
        d = N.f64(dmg) if dmg is not None else None
        self._check(N.lib().ifea_step_set_damage(self._h, N.ptr(d, C.c_double)), "set_damage")

    def checkpoint(self):
        """Forging3DSolver.Checkpoint (:1891-1893): the state at the end of the last increment (see formats)."""
        from .formats import Forging3DCheckpoint
        er = self.element_results()
        m = self.mesh
        return Forging3DCheckpoint(
            tets=m.tets.copy(), on_x0=np.asarray(m.onX0, bool).copy(), on_y0=np.asarray(m.onY0, bool).copy(),
            x=self.x.copy(), vg=self.vg.copy(), vg_prev=None if self.vg_prev is None else self.vg_prev.copy(),
            contact=self.contact.copy(), eps=self.eps.copy(), damage=er["dmg"], sigma=self.el_sig.copy(),
            rate=self.el_rate.copy(), mean=er["elMean"], gap=float(self.gap), ds=float(self.ds), time=float(self.time),
            h=float(self.cfg.h), step=int(self.step_index))

    def restore_from(self, cp, warm: bool = False):
        """RestoreFrom (:1899-1921): positions, velocities, contact, element history, gap, increment and time. Like the
        reference the next increment starts cold (Picard first, no trend history). A checkpoint of another mesh (or any
        checkpoint when cfg.remesh is set) is remeshed at this solver's h and remapped, as the reference always does;
        one of this solver's own mesh with remeshing off is restored onto it as it is, and then `warm` may keep its
        trend history, which replays a run's continuation."""
        m = self.mesh
        same = np.array_equal(cp.tets, m.tets) and np.array_equal(cp.on_x0, m.onX0) and np.array_equal(cp.on_y0, m.onY0)
        if not same or self.cfg.remesh:
            self._setup_mesh(cp.x, cp.tets, cp.on_x0, cp.on_y0)
            self.vg = np.array(cp.vg, dtype=np.float64)
            self.contact = np.array(cp.contact, dtype=np.int32)
            self.set_element_state(cp.eps, cp.sigma)
            self.el_rate = np.array(cp.rate, dtype=np.float64)
            self.set_damage(cp.damage)
            self.gap, self.time, self.step_index = float(cp.gap), float(cp.time), int(cp.step)
            self.ds = min(float(cp.ds) * self.cfg.h / float(cp.h), self.cfg.stroke_increment)
            self.vg_prev, self.warm, self.n_converged = None, False, 1
            if not self.remesh(f"restart from step {cp.step} at h {self.cfg.h:.2f} mm (was {cp.h:.2f})"):
                raise MeshDistortionError(f"Could not remesh the checkpoint of step {cp.step} at h {self.cfg.h:.2f} mm.")
            return
        self.x = np.array(cp.x, dtype=np.float64)
        self.vg = np.array(cp.vg, dtype=np.float64)
        self.contact = np.array(cp.contact, dtype=np.int32)
        self.set_element_state(cp.eps, cp.sigma)
        self.el_rate = np.array(cp.rate, dtype=np.float64)
        self.set_damage(cp.damage)
        self.release = np.zeros(self.nn, dtype=bool)
        self.gap, self.ds, self.time, self.step_index = float(cp.gap), float(cp.ds), float(cp.time), int(cp.step)
        if warm and cp.vg_prev is not None:
            self.vg_prev, self.warm, self.n_converged = np.array(cp.vg_prev, dtype=np.float64), True, 2
        else:
            self.vg_prev, self.warm, self.n_converged = None, bool(warm), 1

    def get_v(self):
        v = np.zeros(self.ndof)
        self._check(N.lib().ifea_step_get_v(self._h, N.ptr(v, C.c_double)), "get_v")
        return v

    def set_v(self, v):
        vv = N.f64(v)
        self._check(N.lib().ifea_step_set_v(self._h, N.ptr(vv, C.c_double)), "set_v")

    def set_element_state(self, eps=None, sigma=None):
        if eps is not None:
            self.eps = np.array(eps, dtype=np.float64)
        if sigma is not None:
            self.el_sig = np.array(sigma, dtype=np.float64)
        self._state_dirty = True

    def info(self, reset: bool = False):
        out = np.zeros(24)
        self._check(N.lib().ifea_step_info(self._h, int(reset), N.ptr(out, C.c_double)), "info")
        return dict(nnz_l=int(out[0]), flops=out[1], supernodes=int(out[2]), device_bytes=int(out[3]),
                    factor_ms=out[4], l_words=int(out[5]), u_words=int(out[6]), levels=int(out[7]),
                    syrk_launches=int(out[8]), syrk_ms=out[9], syrk_flops=out[10], sweeps=int(out[11]),
                    sweep_ms=out[12], sweep_bytes=out[13], factorizations=int(out[14]), factorization_ms=out[15],
                    outer_launches=int(out[16]), outer_ms=out[17], outer_flops=out[18],
                    fp32_factorizations=int(out[19]), fp32_sweeps=int(out[20]), outer_launches_fp32=int(out[21]),
                    outer_ms_fp32=out[22], outer_flops_fp32=out[23])

    def close(self):
        if getattr(self, "_h", None):
            N.lib().ifea_step_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
