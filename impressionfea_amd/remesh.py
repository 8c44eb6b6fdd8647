"""Remesh step of the forging solver: a new mesh of the deformed workpiece and its state remapped onto it.

Restates Forging3DSolver.RemeshTrigger (Forging3DSolver.cs:1867-1880) and RemeshCore (:1958-2102) with the uniform
stuffing lattice: the point locator and the isosurface-stuffing mesher are the library's host geometry
(include/ifea_remesh.h); contact assignment, volume compensation and the placement repair passes are the per-node
logic of placement.py, evaluated on the host (the new mesh is not on the device yet). Not built: the graded lattice
(GradedMesh), local refinement around failure spots (TetRefiner, RefinementLevels' hot spots) and the fine-finish
switch; with their defaults off the reference's remesh is this one.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import placement as PL
from .mesh import TetMesh

FREE, BOTTOM, TOP = PL.FREE, PL.BOTTOM, PL.TOP


def _check(rc, what):
    if rc:
        raise RuntimeError(f"{what} failed ({rc}): {N.lib().ifea_remesh_last_error().decode()}")


class Locator:
    """TetLocator (TetMesh.cs:115-262) over (x, tets) with a bucket grid of cell size `cell`."""

    def __init__(self, x, tets, cell):
        L = N.lib()
        xs, ts = N.f64(x), N.i32(tets)
        self._h = C.c_void_p()
        _check(L.ifea_locator_create(xs.shape[0], N.ptr(xs, C.c_double), ts.shape[0], N.ptr(ts, C.c_int), float(cell),
                                     C.byref(self._h)), "ifea_locator_create")

    def find(self, p):
        p = N.f64(np.atleast_2d(p))
        tet = np.zeros(p.shape[0], np.int32)
        w = np.zeros((p.shape[0], 4))
        _check(N.lib().ifea_locator_find(self._h, p.shape[0], N.ptr(p, C.c_double), N.ptr(tet, C.c_int),
                                         N.ptr(w, C.c_double)), "ifea_locator_find")
        return tet, w

    def find_nearest(self, p):
        p = N.f64(np.atleast_2d(p))
        tet = np.zeros(p.shape[0], np.int32)
        _check(N.lib().ifea_locator_find_nearest(self._h, p.shape[0], N.ptr(p, C.c_double), N.ptr(tet, C.c_int)),
               "ifea_locator_find_nearest")
        return tet

    def interpolate(self, p, field3):
        p, f = N.f64(np.atleast_2d(p)), N.f64(field3)
        out = np.zeros((p.shape[0], 3))
        _check(N.lib().ifea_locator_interpolate(self._h, p.shape[0], N.ptr(p, C.c_double), N.ptr(f, C.c_double),
                                                N.ptr(out, C.c_double)), "ifea_locator_interpolate")
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            N.lib().ifea_locator_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


def stuff(locator: Locator, bmin, bmax, h, symX: bool, symY: bool):
    """IsosurfaceStuffing.Stuff (IsosurfaceStuffing.cs:17-88) of the locator's mesh interior: (points [nn,3], tets
    [ne,4])."""
    L = N.lib()
    lo, hi = N.f64(np.asarray(bmin, float)), N.f64(np.asarray(bmax, float))
    s = C.c_void_p()
    _check(L.ifea_stuff(locator._h, N.ptr(lo, C.c_double), N.ptr(hi, C.c_double), float(h), int(symX), int(symY),
                        C.byref(s)), "ifea_stuff")
    try:
        nn, ne = C.c_int(0), C.c_int(0)
        _check(L.ifea_stuffing_sizes(s, C.byref(nn), C.byref(ne)), "ifea_stuffing_sizes")
        x = np.zeros((nn.value, 3))
        t = np.zeros((ne.value, 4), np.int32)
        if ne.value:
            _check(L.ifea_stuffing_get(s, N.ptr(x, C.c_double), N.ptr(t, C.c_int)), "ifea_stuffing_get")
        return x, t
    finally:
        L.ifea_stuffing_destroy(s)


@dataclass
class RemeshResult:
    x: np.ndarray
    tets: np.ndarray
    onX0: np.ndarray
    onY0: np.ndarray
    contact: np.ndarray
    eps: np.ndarray
    dmg: np.ndarray
    sig: np.ndarray
    rate: np.ndarray
    mean: np.ndarray
    vg: np.ndarray
    stats: dict = field(default_factory=dict)


def remesh_trigger(mesh, x, remesh_quality, h):
    """RemeshTrigger (:1867-1880) without the graded-mesh case: interior smoothing of the current mesh first (x moves
    in place), then a reason when the worst tet is still below remesh_quality or a surface edge is longer than 2.8 h,
    else None."""
    bad = PL.host_bad_tets(mesh)
    if float(PL.tet_quality(x, mesh.tets).min()) < remesh_quality:
        PL.smooth_interior(mesh, x, remesh_quality, 6, bad)
    q = float(PL.tet_quality(x, mesh.tets).min())
    if q < remesh_quality:
        return f"element quality {q:.3f}"
    f = mesh.boundary_faces
    a, b, c = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    el = lambda u, v: np.sqrt(PL.dot3(u - v, u - v))
    max_edge = float(np.max(np.maximum(el(b, a), np.maximum(el(c, b), el(a, c))))) if f.size else 0.0
    return f"surface stretched ({max_edge:.2f} mm edges)" if max_edge > 2.8 * h else None


def _reverts(mesh, xs, stuffed, contact, cfg, gap, h_loc):
    """RemeshCore's revert sweeps (:2039-2057): every node of a tet below 0.03 that is not at its stuffed position goes
    back there (and is released) unless that position is inside a die by more than 0.05 of its local size. Tets in
    ascending order at the current positions, up to four sweeps."""
    import heapq
    t = mesh.tets
    clear_s = np.minimum(cfg.bottom.clearance(stuffed, 0.0), cfg.top.clearance(stuffed, gap))
    total = 0
    for _ in range(4):
        reverted = 0
        heap = [int(e) for e in np.nonzero(PL.tet_quality(xs, t) < 0.03)[0]]
        heapq.heapify(heap)
        queued = set(heap)
        while heap:
            e = heapq.heappop(heap)
            if float(PL.tet_quality(xs, t[e:e + 1])[0]) >= 0.03:
                continue
            for a in range(4):
                n = int(t[e, a])
                if np.array_equal(xs[n], stuffed[n]):
                    continue
                if clear_s[n] < -0.05 * h_loc[n]:
                    continue
                xs[n] = stuffed[n]
                contact[n] = FREE
                reverted += 1
                for e2 in PL.tets_of(mesh, n):
                    if e2 > e and int(e2) not in queued:
                        queued.add(int(e2))
                        heapq.heappush(heap, int(e2))
        total += reverted
        if reverted == 0:
            break
    return total


def remesh_core(cfg, mesh, x, state, gap, vol0, sym_factor, sig_ref, symX, symY, recovery_pass=0):
    """RemeshCore (:1958-2102) on the host: the new mesh with its contact sides and positions, and the element history
    (eps, dmg, sig, rate, mean: nearest old tet of each new centroid) and nodal velocities (interpolated), or None when
    the reference would reject it (stuffing failure, a non-manifold mesh, or quality below 0.02 and not clearly
    better than the old mesh). state: dict of eps, dmg, sig, rate, mean (per old tet) and vg (per old node)."""
    h = cfg.h
    q_old = float(PL.tet_quality(x, mesh.tets).min())
    loc = Locator(x, mesh.tets, 2 * h)
    try:
        xs, ts = stuff(loc, x.min(axis=0), x.max(axis=0), h, symX, symY)
        if ts.shape[0] == 0:
            return None
        onX0 = (xs[:, 0] == 0) & bool(symX)
        onY0 = (xs[:, 1] == 0) & bool(symY)
        try:
            m = TetMesh(xs.shape[0], ts, onX0, onY0)
        except ValueError:
            return None
        stuffed = xs.copy()
        contact = np.zeros(m.nn, np.int32)
        nrm = PL.surface_normals(m, xs)
        h_loc = PL.local_lengths(m, xs, h)
        surf = m.is_surface_node
        cb, ctp = cfg.bottom.clearance(xs, 0.0), cfg.top.clearance(xs, gap)
        tol_c = 0.12 * h_loc
        bot = surf & (cb < tol_c) & (cb <= ctp)
        top = surf & ~bot & (ctp < tol_c)
        contact[bot], contact[top] = BOTTOM, TOP
        for side, die, off in ((BOTTOM, cfg.bottom, 0.0), (TOP, cfg.top, gap)):
            sel = np.nonzero(contact == side)[0]
            if sel.size:
                q, _ = die.project(xs[sel], off)
                q[onX0[sel], 0] = 0.0
                q[onY0[sel], 1] = 0.0
                xs[sel] = q
        PL.compensate_volume(m, xs, contact, vol0, sym_factor, h, cfg, gap)
        bad = PL.host_bad_tets(m)
        n_improve = PL.improve_placement(m, xs, contact, stuffed, nrm, gap, 0.1, h_loc, cfg, bad, recovery_pass)
        n_revert = _reverts(m, xs, stuffed, contact, cfg, gap, h_loc)
        n_smooth = PL.smooth_interior(m, xs, 0.2, 10, bad)
        n_untangle = PL.untangle_surface(m, xs, contact, stuffed, 0.03, 6, cfg, gap, bad)
        n_smooth += PL.smooth_interior(m, xs, 0.2, 4, bad)
        q_new = float(PL.tet_quality(xs, m.tets).min())
        stats = dict(q_old=q_old, q_new=q_new, improve=n_improve, reverts=n_revert, smooth=n_smooth,
                     untangle=n_untangle, nn=m.nn, ne=m.ne)
        if q_new < 0.02 and not (q_new > 0.002 and q_new > q_old + 0.002):
            return None
        tt = m.tets
        cen = 0.25 * (((xs[tt[:, 0]] + xs[tt[:, 1]]) + xs[tt[:, 2]]) + xs[tt[:, 3]])
        found = loc.find_nearest(cen)
        ok = found >= 0
        fi = np.where(ok, found, 0)

        def take(a, default):
            return np.where(ok, np.asarray(a)[fi], default)
        vg = loc.interpolate(xs, state["vg"])
        return RemeshResult(xs, tt, onX0, onY0, contact, take(state["eps"], 0.0), take(state["dmg"], 0.0),
                            take(state["sig"], sig_ref), take(state["rate"], 0.0), take(state["mean"], 0.0), vg,
                            stats)
    finally:
        loc.close()
