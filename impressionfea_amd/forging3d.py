"""Host driver of the MI355X solve step: the increment loop of the reference's Forging3DSolver on a fixed mesh.

Mirrors src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:
  constructor constants and initial contact/velocity   :125-187
  Step (one stroke increment, no remeshing)             :540-638
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


def tet_quality(x, tets):
    """TetMesh.Quality (TetMesh.cs:107-112) of the given tets in the reference's operation order: 6 sqrt(2) V / l2^1.5
    with l2 the mean squared edge length (1 regular, <= 0 inverted, -1 degenerate)."""
    a, b, c, d = (x[tets[..., k]] for k in range(4))
    l2 = (_dot3(b - a, b - a) + _dot3(c - a, c - a) + _dot3(d - a, d - a) + _dot3(c - b, c - b) + _dot3(d - b, d - b)
          + _dot3(d - c, d - c)) / 6.0
    vol = _dot3(b - a, np.cross(c - a, d - a)) / 6.0
    with np.errstate(divide="ignore", invalid="ignore"):
        q = 6 * np.sqrt(2.0) * vol / np.power(l2, 1.5)
    return np.where(l2 > 0, q, -1.0)


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
    """One billet on a fixed mesh, stepped by increments on the GPU (Forging3DSolver.Step without remeshing)."""

    def __init__(self, cfg: StepConfig, x, tets, onX0, onY0, gap0: float, precision: int = 0, perm=None,
                 warm_start=None, first_solve_policy: bool = False, device_frames: bool = True):
        self.cfg = cfg
        self.x = np.array(x, dtype=np.float64)
        self.mesh = TetMesh(self.x.shape[0], tets, np.asarray(onX0, bool), np.asarray(onY0, bool))
        m = self.mesh
        self.nn, self.ne, self.ndof = m.nn, m.ne, 3 * m.nn
        self.nbr_start, self.nbr, self.colPtr, self.rowIdx = build_pattern(m)
        self.nnz = int(self.colPtr[-1])
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
        # LocalLengths (:442-466)
        t = m.tets
        sums = np.zeros(self.nn)
        cnt = np.zeros(self.nn)
        for a in range(4):
            for b in range(a + 1, 4):
                ln = np.linalg.norm(self.x[t[:, a]] - self.x[t[:, b]], axis=1)
                np.add.at(sums, t[:, a], ln)
                np.add.at(sums, t[:, b], ln)
                np.add.at(cnt, t[:, a], 1)
                np.add.at(cnt, t[:, b], 1)
        self.h_local = np.where(cnt > 0, np.minimum(sums / np.maximum(cnt, 1), cfg.h), cfg.h)
        # device step
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
        self.precision = precision
        self.set_params()
        if perm is not None:
            pm = N.i32(perm)
            self._check(L.ifea_step_set_perm(self._h, N.ptr(pm, C.c_int)), "set_perm")
        if first_solve_policy:
            self._check(L.ifea_step_set_option(self._h, N.IFEA_STEP_OPT_FIRST_SOLVE, 1.0), "set_option")
        # PrepareStep's frames on the device (the host's compute_frames is the same arithmetic, kept for checks)
        self.device_frames = bool(device_frames)
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
        p.newton_blend, p.tolerance, p.pcg_tol = c.newton_blend, c.tolerance, 1e-7
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
        from .mlp import trend_start
        guess = self.warm_start(self)
        self.prepare(guess)
        cands = {"learned": guess, "previous": self.vg, "trend": trend_start(self)}
        best, best_e = None, np.inf
        for name, vg in cands.items():
            self.load_guess(vg)
            e = self.assemble(None, matrix=False, newton=False)[0]
            if e < best_e:
                best, best_e = name, e
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
    def surface_normals(self, x):
        """SurfaceNormals (:1540-1562): area-weighted outward unit normals at the surface nodes (zero elsewhere), the
        faces' contributions added in face order."""
        m = self.mesh
        f = m.boundary_faces[~m.face_is_symmetry]
        an = np.cross(x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]])
        # bincount adds its weights in input order: each node sums its faces' contributions in face order
        idx = f.ravel()
        nr = np.stack([np.bincount(idx, weights=np.repeat(an[:, c], 3), minlength=self.nn) for c in range(3)], axis=1)
        nr[m.onX0, 0] = 0.0
        nr[m.onY0, 1] = 0.0
        l2 = _dot3(nr, nr)
        ok = l2 > 0
        nr[ok] = nr[ok] / np.sqrt(l2[ok])[:, None]
        return nr

    def _tets_of(self, n):
        m = self.mesh
        return m.node_tet_list[m.node_tet_start[n]:m.node_tet_start[n + 1]]

    def _worst_around(self, x, n):
        return float(tet_quality(x, self.mesh.tets[self._tets_of(n)]).min())

    def _push_out(self, die, off, p, outward):
        """PushOut (:1570-1592): back to the die surface along the inward normal by bisection within 2 h, else the
        closest point."""
        if _dot3(outward, outward) == 0 or die.clearance(p, off) >= 0:
            return die.project(p, off)[0]
        d = -outward
        h = self.cfg.h
        step, lo, hi = 0.1 * h, 0.0, -1.0
        t = step
        while t <= 2 * h + 1e-12:
            if die.clearance(p + t * d, off) >= 0:
                hi = t
                break
            lo = t
            t += step
        if hi < 0:
            return die.project(p, off)[0]
        for _ in range(25):
            mid = 0.5 * (lo + hi)
            if die.clearance(p + mid * d, off) >= 0:
                hi = mid
            else:
                lo = mid
        return p + hi * d

    def _sym1(self, p, n):
        m = self.mesh
        return np.array([0.0 if m.onX0[n] else p[0], 0.0 if m.onY0[n] else p[1], p[2]])

    def _bad_tets(self, x, q_min):
        """Tets with TetMesh.Quality < q_min at x: the device evaluates every tet, the host re-evaluates those near the
        threshold with the reference's arithmetic so the decision is the reference's."""
        q = np.zeros(self.ne)
        self.first_collapsed(x, q_min, q)
        cand = np.nonzero(q < q_min + 1e-9)[0]
        if cand.size == 0:
            return cand
        return cand[tet_quality(x, self.mesh.tets[cand]) < q_min]

    def improve_placement(self, x, contact, pred, nrm, gap, q_min=0.05):
        """ImprovePlacement (:1625-1679) at the normal pass's tolerances: every die-contact node of a tet below q_min
        tries the closest die point, the exit along the inward normal and, when it barely penetrates, partial
        pushes, keeping whichever most raises the worst quality of its tets (tets visited in ascending order, each
        judged at the current positions). Returns the number of nodes moved; x and contact change in place."""
        import heapq
        cfg, m, t = self.cfg, self.mesh, self.mesh.tets
        total = 0
        for _ in range(4):
            count = 0
            heap = [int(e) for e in self._bad_tets(x, q_min)]
            heapq.heapify(heap)
            queued = set(heap)
            while heap:
                e = heapq.heappop(heap)
                if float(tet_quality(x, t[e:e + 1])[0]) >= q_min:
                    continue
                for a in range(4):
                    n = int(t[e, a])
                    if contact[n] == FREE:
                        continue
                    die, off = (cfg.top, gap) if contact[n] == TOP else (cfg.bottom, 0.0)
                    closest = self._sym1(die.project(pred[n], off)[0], n)
                    cand = [closest]
                    push = self._sym1(self._push_out(die, off, pred[n], nrm[n]), n)
                    if np.sqrt(_dot3(push - pred[n], push - pred[n])) <= 0.5 * self.h_local[n]:
                        cand.append(push)
                    pen = -float(die.clearance(pred[n], off))
                    for f in (0.0, 0.25, 0.5, 0.75):
                        if pen * (1 - f) <= 0.05 * self.h_local[n]:
                            cand.append(pred[n] + f * (closest - pred[n]))
                    old = x[n].copy()
                    best = self._worst_around(x, n)
                    pick = -1
                    for ci, cpt in enumerate(cand):
                        if np.array_equal(cpt, old):
                            continue
                        x[n] = cpt
                        q = self._worst_around(x, n)
                        if q > best + 1e-6:
                            best, pick = q, ci
                    x[n] = cand[pick] if pick >= 0 else old
                    if pick < 0:
                        continue
                    if np.array_equal(cand[pick], pred[n]) and pen < 0:
                        contact[n] = FREE
                    count += 1
                    for e2 in self._tets_of(n):       # tets after e whose quality this move may have changed
                        if e2 > e and int(e2) not in queued:
                            queued.add(int(e2))
                            heapq.heappush(heap, int(e2))
            total += count
            if count == 0:
                break
        return total

    def smooth_interior(self, x, q_target=0.05, sweeps=6):
        """SmoothInterior (:1776-1838): each interior node of a tet below q_target moves towards its neighbours'
        centroid (full, half or quarter step) when that raises the worst quality of its tets; nodes in ascending
        order at the current positions; symmetry-plane nodes stay on their planes. Returns the number of moves."""
        m, t = self.mesh, self.mesh.tets
        moves = 0
        for _ in range(sweeps):
            bad_t = self._bad_tets(x, q_target)
            if bad_t.size == 0:
                break
            nodes = np.unique(t[bad_t].ravel())
            nodes = nodes[~m.is_surface_node[nodes]]
            sweep_moves = 0
            for n in nodes:
                n = int(n)
                te = self._tets_of(n)
                tt = t[te]
                others = tt[tt != n]
                if others.size == 0:
                    continue
                s = np.zeros(3)
                for q in others:                    # the reference's running sum, in its order
                    s = s + x[q]
                worst = float(tet_quality(x, tt).min())
                old = x[n].copy()
                c = (1.0 / others.size) * s
                for lam in (1.0, 0.5, 0.25):
                    p = old + lam * (c - old)
                    x[n] = self._sym1(p, n)
                    if float(tet_quality(x, tt).min()) > worst + 1e-6:
                        sweep_moves += 1
                        break
                    x[n] = old
            moves += sweep_moves
            if sweep_moves == 0:
                break
        return moves

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
        self.time += dt
        return ""

    def step(self) -> IncrementResult:
        """One stroke increment (Forging3DSolver.Step :540-638) on the fixed mesh."""
        cfg = self.cfg
        self.step_index += 1
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
        why = ""
        for attempt in range(6):
            dta = dt / (1 << attempt)
            why = self.try_update(dta)
            if not why:
                dt = dta
                break
        if why:
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
        return IncrementResult(self.step_index, load, reaction, st["iterations"], bool(st["converged"]), dt, self.gap,
                               st, 1e3 * ((t1 - t0) + (t3 - t2)), 1e3 * (t2 - t1))

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
        """RestoreFrom (:1899-1921) on this solver's own mesh: positions, velocities, contact, element history, gap,
        increment and time. Like the reference the next increment starts cold (Picard first, no trend history) unless
        `warm` keeps the checkpoint's trend history, which replays a run's continuation. A checkpoint of another mesh
        needs the reference's remesh and remap, which this solver does not do."""
        m = self.mesh
        if not (np.array_equal(cp.tets, m.tets) and np.array_equal(cp.on_x0, m.onX0) and np.array_equal(cp.on_y0, m.onY0)):
            raise NotImplementedError("the checkpoint holds another mesh: restoring it needs a remesh and remap")
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
