"""Node placement on the workpiece surface and mesh repair, shared by TryUpdate and the remesh step.

Restates, per node and in the reference's order of operations, Forging3DSolver.cs:
  SurfaceNormals    :1540-1562      PushOut            :1570-1592
  ImprovePlacement  :1625-1679      UntangleSurface    :1698-1772
  SmoothInterior    :1776-1838      CompensateVolume   :2108-2145
  LocalLengths      :443-466        TetMesh.Quality    TetMesh.cs:107-112
The passes visit only the tets below a quality threshold: a `bad_tets(x, q)` callable lists them (the device's quality
kernel during an increment, a host evaluation for a mesh that is not on the device yet). Positions and contact sides
change in place; every function returns its number of node moves.
"""
from __future__ import annotations

import heapq

import numpy as np

FREE, BOTTOM, TOP = 0, 1, 2
# CornerPenetrationLadder (:1599): the penetration a node may keep at a die corner, times its local size, in the normal
# pass and the first recovery pass (the lenient passes with refinement are not built)
CORNER_LADDER = (0.05, 0.1)


def dot3(a, b):
    """Row-wise a . b as (a0 b0 + a1 b1) + a2 b2: the reference's Vec3.Dot order, never a fused or reassociated sum."""
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def tet_quality(x, tets):
    """TetMesh.Quality (TetMesh.cs:107-112) of the given tets in the reference's operation order: 6 sqrt(2) V / l2^1.5
    with l2 the mean squared edge length (1 regular, <= 0 inverted, -1 degenerate)."""
    return tet_quality_pts(x[tets])


def tet_quality_pts(P):
    """tet_quality of tets given by their corner positions P[..., 4, 3] (the same operations, element by element)."""
    a, b, c, d = (P[..., k, :] for k in range(4))
    l2 = (dot3(b - a, b - a) + dot3(c - a, c - a) + dot3(d - a, d - a) + dot3(c - b, c - b) + dot3(d - b, d - b)
          + dot3(d - c, d - c)) / 6.0
    vol = dot3(b - a, np.cross(c - a, d - a)) / 6.0
    with np.errstate(divide="ignore", invalid="ignore"):
        q = 6 * np.sqrt(2.0) * vol / np.power(l2, 1.5)
    return np.where(l2 > 0, q, -1.0)


def tet_volumes(x, tets):
    """TetMesh.Volume (TetMesh.cs:104): (b - a) . ((c - a) x (d - a)) / 6."""
    a, b, c, d = (x[tets[..., k]] for k in range(4))
    return dot3(b - a, np.cross(c - a, d - a)) / 6.0


def host_bad_tets(mesh):
    """bad_tets for a mesh evaluated on the host."""
    def bad(x, q_min):
        return np.nonzero(tet_quality(x, mesh.tets) < q_min)[0]
    return bad


def local_lengths(mesh, x, h):
    """LocalLengths (:443-466): the mean length of each node's edges (per tet, every edge counted), capped at h; h for
    nodes without tets. The sums run tet by tet, edge by edge, as in the reference."""
    t = mesh.tets
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    lens = np.stack([np.sqrt(dot3(x[t[:, a]] - x[t[:, b]], x[t[:, a]] - x[t[:, b]])) for a, b in pairs], axis=1)
    idx = np.stack([np.stack([t[:, a], t[:, b]], axis=1) for a, b in pairs], axis=1)   # (ne, 6, 2): i then j
    w = np.repeat(lens[:, :, None], 2, axis=2)
    s = np.bincount(idx.ravel(), weights=w.ravel(), minlength=mesh.nn)
    cnt = np.bincount(idx.ravel(), minlength=mesh.nn)
    return np.where(cnt > 0, np.minimum(s / np.maximum(cnt, 1), h), h)


def surface_normals(mesh, x):
    """SurfaceNormals (:1540-1562): area-weighted outward unit normals at the surface nodes (zero elsewhere), the faces'
    contributions added in face order."""
    f = mesh.boundary_faces[~mesh.face_is_symmetry]
    an = np.cross(x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]])
    # bincount adds its weights in input order: each node sums its faces' contributions in face order
    idx = f.ravel()
    nr = np.stack([np.bincount(idx, weights=np.repeat(an[:, c], 3), minlength=mesh.nn) for c in range(3)], axis=1)
    nr[mesh.onX0, 0] = 0.0
    nr[mesh.onY0, 1] = 0.0
    l2 = dot3(nr, nr)
    ok = l2 > 0
    nr[ok] = nr[ok] / np.sqrt(l2[ok])[:, None]
    return nr


def tets_of(mesh, n):
    return mesh.node_tet_list[mesh.node_tet_start[n]:mesh.node_tet_start[n + 1]]


def worst_around(mesh, x, n):
    return float(tet_quality(x, mesh.tets[tets_of(mesh, n)]).min())


def worst_around_at(mesh, x, n, positions):
    """worst_around(mesh, x, n) with node n moved to each of positions[k] in turn (one evaluation for all of them)."""
    tt = mesh.tets[tets_of(mesh, n)]
    P = np.repeat(x[tt][None], len(positions), axis=0)       # (k, tets, 4, 3)
    at = tt == n
    P[:, at] = np.asarray(positions, dtype=np.float64)[:, None, :]
    return tet_quality_pts(P).min(axis=1)


def sym1(mesh, p, n):
    return np.array([0.0 if mesh.onX0[n] else p[0], 0.0 if mesh.onY0[n] else p[1], p[2]])


def push_out(die, off, p, outward, h):
    """PushOut (:1570-1592): back to the die surface along the inward normal by bisection within 2 h, else the closest
    point."""
    if dot3(outward, outward) == 0 or die.clearance(p, off) >= 0:
        return die.project(p, off)[0]
    d = -outward
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


def push_out_many(die, off, p, outward, h):
    """push_out for many points at once: the same steps and bisections per point, in lockstep."""
    p = np.atleast_2d(p)
    out = np.array(die.project(p, off)[0], dtype=np.float64)
    nz = dot3(outward, outward) != 0
    act = nz & (die.clearance(p, off) < 0)
    if not act.any():
        return out
    idx = np.nonzero(act)[0]
    pa, d = p[idx], -outward[idx]
    step = 0.1 * h
    lo = np.zeros(idx.size)
    hi = np.full(idx.size, -1.0)
    found = np.zeros(idx.size, bool)
    t = step
    while t <= 2 * h + 1e-12:
        open_ = ~found
        if not open_.any():
            break
        hit = open_ & (die.clearance(pa + t * d, off) >= 0)
        hi[hit] = t
        found |= hit
        lo[open_ & ~hit] = t
        t += step
    for _ in range(25):
        mid = 0.5 * (lo + hi)
        ok = die.clearance(pa + mid[:, None] * d, off) >= 0
        hi = np.where(found & ok, mid, hi)
        lo = np.where(found & ~ok, mid, lo)
    res = pa + hi[:, None] * d
    out[idx[found]] = res[found]
    return out


def die_of(cfg, side, gap):
    return (cfg.top, gap) if side == TOP else (cfg.bottom, 0.0)


def improve_placement(mesh, x, contact, pred, nrm, gap, q_min, h_local, cfg, bad_tets, recovery_pass=0):
    """ImprovePlacement (:1625-1679) with the strict tolerances (Tol = the local size): every die-contact node of a tet
    below q_min tries the closest die point, the exit along the inward normal and, when it barely penetrates, partial
    pushes, keeping whichever most raises the worst quality of its tets. Tets are visited in ascending order, each
    judged at the current positions (a heap holds the bad tets and the later tets of every moved node)."""
    t = mesh.tets
    corner = CORNER_LADDER[min(recovery_pass, len(CORNER_LADDER) - 1)]
    # a node's candidates depend only on its predicted position, normal and die: computed once per die side, for the
    # contact nodes of the tets below q_min in one go, for others when first needed
    cache = {}

    def candidates(n):
        key = (n, int(contact[n]))
        if key not in cache:
            die, off = die_of(cfg, contact[n], gap)
            cache[key] = (die.project(pred[n], off)[0], push_out(die, off, pred[n], nrm[n], cfg.h),
                          -float(die.clearance(pred[n], off)))
        return cache[key]

    first = bad_tets(x, q_min)
    if first.size:
        nodes = np.unique(t[first].ravel())
        for side in (BOTTOM, TOP):
            sel = nodes[np.asarray(contact)[nodes] == side]
            if sel.size:
                die, off = die_of(cfg, side, gap)
                cl = die.project(pred[sel], off)[0]
                pu = push_out_many(die, off, pred[sel], nrm[sel], cfg.h)
                pe = -die.clearance(pred[sel], off)
                for i, n in enumerate(sel):
                    cache[(int(n), side)] = (cl[i], pu[i], float(pe[i]))
    total = 0
    # qualities evaluated at the current positions stay valid until a node moves: the queued tets are evaluated in one
    # call and a node's candidate qualities are kept, both dropped at the first move (moves are rare; most sweeps
    # only confirm that no candidate helps)
    qcache, ccache = {}, {}
    for sweep in range(4):
        count = 0
        heap = [int(e) for e in (first if sweep == 0 else bad_tets(x, q_min))]
        heapq.heapify(heap)
        queued = set(heap)
        if heap:
            qcache = dict(zip(heap, tet_quality(x, t[np.asarray(heap)]).tolist()))
        ccache = {}
        while heap:
            e = heapq.heappop(heap)
            qe = qcache.get(e)
            if qe is None:
                qe = float(tet_quality(x, t[e:e + 1])[0])
            if qe >= q_min:
                continue
            for a in range(4):
                n = int(t[e, a])
                if contact[n] == FREE:
                    continue
                closest, push, pen = candidates(n)
                closest = sym1(mesh, closest, n)
                cand = [closest]
                push = sym1(mesh, push, n)
                if np.sqrt(dot3(push - pred[n], push - pred[n])) <= 0.5 * h_local[n]:
                    cand.append(push)
                for f in (0.0, 0.25, 0.5, 0.75):
                    if pen * (1 - f) <= corner * h_local[n]:
                        cand.append(pred[n] + f * (closest - pred[n]))
                old = x[n].copy()
                # the worst quality around n at its position and at every candidate, evaluated at once; then the
                # reference's choice, candidate by candidate
                qs = ccache.get(n)
                if qs is None:
                    qs = ccache[n] = worst_around_at(mesh, x, n, [old] + cand)
                best = float(qs[0])
                pick = -1
                for ci, cpt in enumerate(cand):
                    if np.array_equal(cpt, old):
                        continue
                    q = float(qs[ci + 1])
                    if q > best + 1e-6:
                        best, pick = q, ci
                x[n] = cand[pick] if pick >= 0 else old
                if pick < 0:
                    continue
                if np.array_equal(cand[pick], pred[n]) and pen < 0:
                    contact[n] = FREE
                count += 1
                qcache, ccache = {}, {}
                for e2 in tets_of(mesh, n):       # tets after e whose quality this move may have changed
                    if e2 > e and int(e2) not in queued:
                        queued.add(int(e2))
                        heapq.heappush(heap, int(e2))
        total += count
        if count == 0:
            break
    return total


def _neighbour_centroid(mesh, x, n):
    tt = mesh.tets[tets_of(mesh, n)]
    others = tt[tt != n]
    if others.size == 0:
        return tt, None
    s = np.zeros(3)
    for q in others:                    # the reference's running sum, in its order
        s = s + x[q]
    return tt, (1.0 / others.size) * s


def smooth_interior(mesh, x, q_target, sweeps, bad_tets=None):
    """SmoothInterior (:1776-1838): each interior node of a tet below q_target moves towards its neighbours' centroid
    (full, half or quarter step) when that raises the worst quality of its tets; nodes in ascending order at the
    current positions; symmetry-plane nodes stay on their planes. The library's host routine (ifea_smooth_interior)
    does the sweeps; bad_tets is not needed (it keeps every tet's quality current itself)."""
    import ctypes as C
    from . import _native as N
    if not x.flags["C_CONTIGUOUS"] or x.dtype != np.float64:
        raise ValueError("x must be a C-contiguous float64 array (it is updated in place)")
    if not hasattr(mesh, "_smooth_args"):
        mesh._smooth_args = (N.i32(mesh.tets), N.i32(mesh.node_tet_start), N.i32(mesh.node_tet_list),
                             np.ascontiguousarray(mesh.is_surface_node, np.uint8),
                             np.ascontiguousarray(mesh.onX0, np.uint8), np.ascontiguousarray(mesh.onY0, np.uint8))
    t, st, lst, su, ox, oy = mesh._smooth_args
    moves = C.c_int(0)
    rc = N.lib().ifea_smooth_interior(mesh.nn, mesh.ne, N.ptr(t, C.c_int), N.ptr(st, C.c_int), N.ptr(lst, C.c_int),
                                      N.ptr(su, C.c_ubyte), N.ptr(ox, C.c_ubyte), N.ptr(oy, C.c_ubyte),
                                      N.ptr(x, C.c_double), float(q_target), int(sweeps), C.byref(moves))
    if rc:
        raise RuntimeError(f"ifea_smooth_interior failed ({rc}): {N.lib().ifea_remesh_last_error().decode()}")
    return moves.value


def untangle_surface(mesh, x, contact, stuffed, q_target, sweeps, cfg, gap, bad_tets):
    """UntangleSurface (:1698-1772): each surface node of a tet below q_target tries its stuffed position and its
    neighbours' centroid at full, half and quarter steps outside both dies, keeping the candidate that most raises the
    worst quality of its tets; a node that moves is released (contact FREE)."""
    t = mesh.tets
    moves = 0
    for _ in range(sweeps):
        bad_t = bad_tets(x, q_target)
        if bad_t.size == 0:
            break
        nodes = np.unique(t[bad_t].ravel())
        nodes = nodes[mesh.is_surface_node[nodes]]
        sweep_moves = 0
        for n in nodes:
            n = int(n)
            tt, centroid = _neighbour_centroid(mesh, x, n)
            if centroid is None:
                continue
            old = x[n].copy()
            best, best_q = old, worst_around(mesh, x, n)
            for target in (stuffed[n], centroid):
                for lam in (1.0, 0.5, 0.25):
                    p = sym1(mesh, old + lam * (target - old), n)
                    if min(float(cfg.bottom.clearance(p, 0.0)), float(cfg.top.clearance(p, gap))) < 0:
                        continue
                    x[n] = p
                    q = worst_around(mesh, x, n)
                    if q > best_q + 1e-6:
                        best, best_q = p, q
            x[n] = best
            if not np.array_equal(best, old):
                contact[n] = FREE
                sweep_moves += 1
        moves += sweep_moves
        if sweep_moves == 0:
            break
    return moves


def compensate_volume(mesh, x, contact, vol0, sym_factor, h, cfg, gap):
    """CompensateVolume (:2108-2145): the free-surface nodes move along their area-weighted normals by one common shift
    (at most h) that restores the initial volume, each only where it stays clear of both dies."""
    vol = float(np.cumsum(tet_volumes(x, mesh.tets))[-1])   # the reference's running sum, tet by tet
    deficit = vol0 / sym_factor - vol
    f = mesh.boundary_faces[~mesh.face_is_symmetry]
    an = 0.5 * np.cross(x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]])
    idx = f.ravel()
    nr = np.stack([np.bincount(idx, weights=np.repeat(an[:, c], 3), minlength=mesh.nn) for c in range(3)], axis=1)
    use = mesh.is_surface_node & (np.asarray(contact) == FREE)
    nr[~use] = 0.0
    nr[use & mesh.onX0, 0] = 0.0
    nr[use & mesh.onY0, 1] = 0.0
    ln = np.sqrt(dot3(nr, nr))
    terms = ln[use] / 3
    rate = float(np.cumsum(terms)[-1]) if terms.size else 0.0   # in node order
    ok = use & (ln > 0)
    nr[ok] = nr[ok] / ln[ok][:, None]
    if rate <= 0:
        return
    shift = min(max(deficit / rate, -h), h)
    moved = x + shift * nr
    keep = ok & (cfg.bottom.clearance(moved, 0.0) > 0) & (cfg.top.clearance(moved, gap) > 0)
    x[keep] = moved[keep]
