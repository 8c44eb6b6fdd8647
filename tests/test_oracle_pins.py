"""Pins of the oracle functions that the assembly-level tests leave open: BuildFriction, FinalEvaluation's element mean
stress, the Kernel's deviatoric stress, Principal.Max3, AccumulateDamage, the oracle's minimum-degree ordering and the
ordering seam against MKL PARDISO's own permutation.

Each check is a closed form, an invariant, a library routine (numpy's symmetric eigensolver) or brute force, chosen so
that a plausible slip (a wrong area share, a dropped 1/sqrt(3), a wrong shear index, a permutation read the wrong way
round) fails it.
"""
import ctypes as C

import numpy as np
import pytest

from impressionfea_amd.config import DieSpec, StepConfig, upset_small
from impressionfea_amd.mesh import TetMesh, build_pattern, kuhn_box, kuhn_cylinder


def _box(n=3, h=2.0, jitter=0.1, seed=0, quarter=False):
    """n^3 Kuhn box of side L = n h with interior jitter only (the faces stay exact planes)."""
    I, J, K = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    x = np.stack([I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()], 1) * h
    x = x.astype(np.float64)
    rng = np.random.default_rng(seed)
    inner = np.all((x > 0) & (x < n * h), axis=1)
    x[inner] += rng.uniform(-jitter, jitter, size=(inner.sum(), 3)) * h
    t = kuhn_box(n, n, n).astype(np.int32)
    if quarter:
        return x, t, x[:, 0] == 0.0, x[:, 1] == 0.0
    z = np.zeros(x.shape[0], bool)
    return x, t, z, z.copy()


def _cfg(L, friction=0.3, **kw):
    return StepConfig(friction=friction, size=L, h=L / 3, bottom=DieSpec(0, 0, (-100.0,) + (0.0,) * 7),
                      top=DieSpec(1, 0, (100.0,) + (0.0,) * 7), **kw)


def _oracle(cfg, x, t, ox, oy):
    import oracle as O
    return O.Oracle(cfg, x, t, ox, oy, 0.0)


# ------------------------------------------------------------------------------------------- BuildFriction (:829-850)
@pytest.mark.parametrize("quarter", [False, True])
def test_friction_area_is_the_free_surface_area(oracle_lib, quarter):
    """sum_n aFric_n = the area of the non-symmetry boundary: 6 L^2 for the whole box, 4 L^2 with the symmetry planes
    x = 0 and y = 0 excluded; interior nodes get no area."""
    L = 6.0
    x, t, ox, oy = _box(3, 2.0, quarter=quarter)
    o = _oracle(_cfg(L), x, t, ox, oy)
    o.prepare()
    g = o.geometry()
    want = (4.0 if quarter else 6.0) * L * L
    assert g["aFric"].sum() == pytest.approx(want, rel=1e-13)
    m = TetMesh(x.shape[0], t, ox, oy)
    assert np.all(g["aFric"][~m.is_surface_node] == 0.0)
    assert np.all(g["aFric"][m.is_surface_node] > 0.0)


def test_friction_shear_stress_is_mean_of_owner_stress_over_sqrt3(oracle_lib):
    """kFric_n = sigma / sqrt(3) under a uniform flow stress; with per-element stresses it is the mean of
    sigma_owner / sqrt(3) over the node's non-symmetry boundary faces (brute force over the face list)."""
    L = 6.0
    x, t, ox, oy = _box(3, 2.0, quarter=True)
    o = _oracle(_cfg(L), x, t, ox, oy)
    s = 41.5
    o.set_state(elSig=np.full(o.ne, s))
    o.prepare()
    k = o.geometry()["kFric"]
    m = TetMesh(x.shape[0], t, ox, oy)
    assert np.abs(k[m.is_surface_node] - s / np.sqrt(3.0)).max() <= 1e-14 * s
    assert np.all(k[~m.is_surface_node] == 0.0)
    sig = np.random.default_rng(2).uniform(10.0, 90.0, o.ne)
    o.set_state(elSig=sig)
    o.prepare()
    k = o.geometry()["kFric"]
    bf, ow, fs, _ = o.topology()
    faces = {n: [] for n in range(o.nn)}
    for f in range(len(ow)):
        if fs[f]:
            continue
        for n in bf[f]:
            faces[int(n)].append(sig[ow[f]])
    for n in range(o.nn):
        want = np.mean(faces[n]) / np.sqrt(3.0) if faces[n] else 0.0
        assert k[n] == pytest.approx(want, rel=1e-14, abs=0.0)


def test_friction_power_at_large_slip(oracle_lib):
    """Every surface node on the flat bottom die, rigid sliding at |s| = U >> u0, uniform flow stress: no strain rate
    and no divergence, so the ram power is the friction power alone, m k A U (2/pi) atan(U/u0) with k = sigma/sqrt(3)
    and A = 6 L^2, i.e. m k A U to within (2/pi) u0 / U; the energy is m k A (2/pi) (U atan z - u0/2 ln(1 + z^2)).
    The penalty is made small: K_vol v of a rigid field vanishes only to rounding, which is of order eps |K_vol| |v|."""
    L = 6.0
    x, t, ox, oy = _box(3, 2.0)
    mfr = 0.3
    o = _oracle(_cfg(L, friction=mfr, penalty_factor=1e-6), x, t, ox, oy)
    s = 50.0
    m = TetMesh(x.shape[0], t, ox, oy)
    contact = np.where(m.is_surface_node, 1, 0).astype(np.int32)
    o.set_state(contact=contact, elSig=np.full(o.ne, s))
    o.prepare()
    fr = o.frames()
    u0 = o.constants()["u0"]
    U = 1e4 * u0
    vg = np.tile([U, 0.0, 0.0], (o.nn, 1))
    v = np.einsum("nij,nj->ni", fr["ax"], vg).ravel()
    assert np.abs(v[fr["fixed"]]).max() <= 1e-12 * U      # the die normal is z: sliding leaves the fixed DOF at 0
    e, pw, r, _ = o.assemble(v, matrix=False, newton=False)
    kA = mfr * s / np.sqrt(3.0) * 6 * L * L
    z = U / u0
    assert pw == pytest.approx(kA * U * (2 / np.pi) * np.arctan(z), rel=1e-12)
    assert pw == pytest.approx(kA * U, rel=(2 / np.pi) / z * 1.01)
    assert e == pytest.approx(kA * (2 / np.pi) * (U * np.arctan(z) - 0.5 * u0 * np.log1p(z * z)), rel=1e-12)
    # the residual on each sliding node's free axes is c atan(z)/|s| * s: along x, summed over the nodes, m k A (2/pi)
    # atan(z) (the local x axis of a flat-die node is the global x or y axis)
    rg = np.einsum("nji,nj->ni", fr["ax"], r.reshape(-1, 3))
    assert rg[:, 0].sum() == pytest.approx(kA * (2 / np.pi) * np.arctan(z), rel=1e-12)
    assert np.abs(rg[:, 1:]).max() <= 1e-9 * np.abs(rg[:, 0]).max()   # rounding of the rigid field's element terms


# -------------------------------------------------------------- Kernel deviatoric stress, FinalEvaluation mean stress
def test_kernel_deviatoric_stress_closed_form(oracle_lib):
    """Homogeneous compression v = (r x/2, r y/2, -r z): the deviatoric stress of every tet is sigma (1/3, 1/3, -2/3)
    with no shear (elDev order xx, yy, zz, xy, yz, zx); on a random field sqrt(3/2 s:s) = sigma wherever the rate is
    above the cut-off (the von Mises invariant of s = mu (2/3 dev, gamma/3))."""
    wl = upset_small(nxy=4, nz=4, friction=0.0)
    x, t, ox, oy = kuhn_cylinder(wl.diameter, wl.height, wl.nxy, wl.nz, seed=2)
    import oracle as O
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.prepare()
    fr = o.frames()
    r = 0.05
    v = np.einsum("nij,nj->ni", fr["ax"], x * np.array([0.5 * r, 0.5 * r, -r])).ravel()
    o.assemble(v, matrix=False, newton=False)
    st = o.element_state()
    sig = st["elSig"][:, None]
    want = np.hstack([sig / 3, sig / 3, -2 * sig / 3, np.zeros((o.ne, 3))])
    assert np.abs(st["elDev"] - want).max() <= 1e-12 * sig.max()
    rng = np.random.default_rng(3)
    v2 = v + 0.02 * rng.standard_normal(v.size)
    o.assemble(v2, matrix=False, newton=False)
    st = o.element_state()
    d = st["elDev"]
    vm = np.sqrt(1.5 * ((d[:, :3] ** 2).sum(1) + 2 * (d[:, 3:] ** 2).sum(1)))
    hot = st["elRate"] > o.constants()["rateCut"]
    assert hot.sum() > 0.9 * o.ne
    assert np.abs(vm[hot] - st["elSig"][hot]).max() <= 1e-12 * st["elSig"].max()


def test_final_mean_stress_linear_field(oracle_lib):
    """FinalEvaluation (:1440-1450): for v = (a x, b y, c z) every nodal volumetric rate g_n . v is a + b + c, so every
    element mean stress is kPen (a + b + c)."""
    L = 6.0
    x, t, ox, oy = _box(3, 2.0, jitter=0.15, seed=4)
    o = _oracle(_cfg(L, friction=0.0, penalty_factor=1e3), x, t, ox, oy)
    o.prepare()
    fr = o.frames()
    a, b, c = 0.3, -0.7, 0.15
    o.set_v(np.einsum("nij,nj->ni", fr["ax"], x * np.array([a, b, c])).ravel())
    o.final()
    mean = o.element_state()["elMean"]
    kpen = o.constants()["kPen"]
    assert np.abs(mean - kpen * (a + b + c)).max() <= 1e-11 * kpen


# --------------------------------------------------------------------------------------- Principal.Max3 (:7-19)
def _rot(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return q


def _t6(T):
    return np.array([T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[1, 2], T[2, 0]])


def test_max3_diagonal_rotated_random_repeated(oracle_lib):
    import oracle as O
    rng = np.random.default_rng(5)
    # diagonal tensors take the early exit
    d = np.array([[3.0, -1.0, 2.0], [-5.0, -2.0, -9.0], [0.0, 0.0, 0.0], [1e-8, 4e-8, -2e-8]])
    t6 = np.hstack([d, np.zeros((4, 3))])
    assert np.array_equal(O.max3(t6), d.max(axis=1))
    # rotated diagonal tensors, repeated eigenvalues included (the acos clamp), against the eigenvalues put in
    lams = [(3.0, -1.0, 2.0), (2.0, -1.0, -1.0), (1.0, 1.0, 3.0), (-4.0, -4.0, -4.0 + 1e-9), (250.0, -80.0, 0.5)]
    for lam in lams:
        Q = _rot(rng)
        T = Q @ np.diag(lam) @ Q.T
        got = O.max3(_t6(T))[0]
        assert got == pytest.approx(max(lam), abs=1e-12 * max(abs(v) for v in lam))
    # random symmetric tensors against LAPACK's symmetric eigensolver; the trace and the largest eigenvalue's
    # characteristic-polynomial residual as further invariants
    T = rng.standard_normal((200, 3, 3)) * rng.uniform(0.1, 100.0, (200, 1, 1))
    T = 0.5 * (T + T.transpose(0, 2, 1))
    got = O.max3(np.stack([_t6(a) for a in T]))
    w = np.linalg.eigvalsh(T)
    scale = np.abs(w).max(axis=1)
    assert np.all(np.abs(got - w[:, -1]) <= 1e-12 * scale)
    det = np.linalg.det(T - got[:, None, None] * np.eye(3))
    assert np.all(np.abs(det) <= 1e-9 * scale ** 3)


# -------------------------------------------------------------------------------------- AccumulateDamage (:1844-1855)
def test_damage_closed_forms(oracle_lib):
    """Cockcroft-Latham D += max(sigma_1, 0) / sigma_bar * rate * dt: uniaxial tension along any axis gives
    rate dt, pure shear rate dt / sqrt(3), uniaxial compression nothing; tets at or below the rate cut-off, or with
    no flow stress, take no damage; damage accumulates over calls."""
    L = 6.0
    x, t, ox, oy = _box(2, 3.0)
    o = _oracle(_cfg(L, friction=0.0), x, t, ox, oy)
    ne = o.ne
    rng = np.random.default_rng(6)
    cut = o.constants()["rateCut"]
    sbar = rng.uniform(20.0, 80.0, ne)
    rate = rng.uniform(2.0, 10.0, ne) * cut
    kind = np.arange(ne) % 5          # 0 tension, 1 shear, 2 compression, 3 tension below the cut-off, 4 no stress
    dev, mean = np.zeros((ne, 6)), np.zeros(ne)
    for e in range(ne):
        Q = _rot(rng)
        n = Q[:, 0]
        if kind[e] == 1:
            tau = sbar[e] / np.sqrt(3.0)
            S = tau * (np.outer(Q[:, 0], Q[:, 1]) + np.outer(Q[:, 1], Q[:, 0]))
        elif kind[e] == 2:
            S = -sbar[e] * np.outer(n, n)
        else:
            S = sbar[e] * np.outer(n, n)
        m = np.trace(S) / 3
        dev[e] = _t6(S - m * np.eye(3))
        mean[e] = m
    rate[kind == 3] = cut
    sbar[kind == 4] = 0.0
    o.set_element_state(elSig=sbar, elRate=rate, elDev=dev.ravel(), elMean=mean)
    dt = 0.05
    o.accumulate_damage(dt)
    D = o.element_state()["dmg"]
    want = np.select([kind == 0, kind == 1], [rate * dt, rate * dt / np.sqrt(3.0)], 0.0)
    # tension and shear to rounding; under uniaxial compression the largest eigenvalue (0) is a double root, where the
    # trigonometric solution is only accurate to sqrt(eps) (acos at -1), so a residue of ~1e-8 of rate dt may remain
    tol = np.where(kind == 2, 1e-7, 1e-12) * rate * dt
    assert np.all(np.abs(D - want) <= tol)
    assert np.all(D[kind >= 3] == 0.0)
    o.accumulate_damage(2 * dt)
    assert np.all(np.abs(o.element_state()["dmg"] - 3 * want) <= 3 * tol)


# ---------------------------------------------------------------------------------- the oracle's own ordering
def _min_degree_brute(nn, ns, nb, order):
    """Replays an elimination order on Python sets and checks the minimum-degree property at every step: the node
    taken has the smallest weighted degree (3 DOFs per neighbour) among the nodes left, and the lowest index among
    ties; eliminating it joins its neighbours into a clique."""
    adj = [set(int(q) for q in nb[ns[v]:ns[v + 1]] if q != v) for v in range(nn)]
    left = set(range(nn))
    for k, v in enumerate(order):
        deg = {u: 3 * len(adj[u]) for u in left}
        dmin = min(deg.values())
        assert deg[v] == dmin, (k, v, deg[v], dmin)
        assert v == min(u for u in left if deg[u] == dmin), (k, v)
        nbv = adj[v]
        for u in nbv:
            adj[u] |= nbv
            adj[u].discard(u)
            adj[u].discard(v)
        left.discard(v)
        adj[v] = set()
    assert not left


@pytest.mark.parametrize("nxy,nz,seed", [(2, 2, 0), (3, 3, 1)])
def test_node_min_degree_property(oracle_lib, nxy, nz, seed):
    """The oracle's NodeMinDegree ordering is an exact minimum-degree order on the node graph (brute force), with the
    three DOFs of a node eliminated together."""
    import oracle as O
    x, t, ox, oy = kuhn_cylinder(40, 60, nxy, nz, seed=seed)
    m = TetMesh(x.shape[0], t, ox, oy)
    ns, nb, _, _ = build_pattern(m)
    perm = O.min_degree_nodes(m.nn, ns, nb)
    order = perm[0::3] // 3
    assert np.array_equal(perm.reshape(-1, 3), 3 * order[:, None] + np.arange(3))
    assert sorted(order) == list(range(m.nn))
    _min_degree_brute(m.nn, ns, nb, [int(v) for v in order])


def test_min_degree_brute_force_catches_a_wrong_order(oracle_lib):
    """The brute-force check above rejects an order that is not minimum degree (the oracle's order with two steps
    swapped), so it pins something."""
    import oracle as O
    x, t, ox, oy = kuhn_cylinder(40, 60, 2, 2, seed=0)
    m = TetMesh(x.shape[0], t, ox, oy)
    ns, nb, _, _ = build_pattern(m)
    order = [int(v) for v in O.min_degree_nodes(m.nn, ns, nb)[0::3] // 3]
    order[0], order[-1] = order[-1], order[0]
    with pytest.raises(AssertionError):
        _min_degree_brute(m.nn, ns, nb, order)


# ------------------------------------------------------------------- the ordering seam against MKL PARDISO (:20-31)
def test_mkl_returned_permutation_through_the_symbolic_analysis(oracle_lib, c1_mesh):
    """PARDISO's own METIS ordering of the configured upset system (iparm[4] = 2 returns it) fed to this library's
    symbolic analysis as a caller ordering (the ifea_spd_analyze_perm route): the ordering is kept as given, the
    elimination tree and column counts equal brute-force elimination bit for bit, and MKL's reported factor size
    (iparm[17]) is the exact structural count plus MKL's relaxed-supernode padding (0-5 %). The same holds for the
    identity ordering, where nothing but the padding can differ. MKL's parallel METIS is not deterministic run to run,
    so its permutation is taken once and compared as data."""
    import oracle as O
    from impressionfea_amd import _native as N
    if not O.pardiso_available():
        pytest.skip("MKL runtime not present")
    wl, x, t, ox, oy = c1_mesh
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.initial(0.0)
    o.prepare()
    fr = o.frames()
    _, _, _, val = o.assemble(o.v(), True, True)
    ns, nb, cp, ri = o.pattern()
    n = o.ndof
    fixed = fr["fixed"]
    col = np.repeat(np.arange(n), np.diff(cp))
    A = val.copy()
    msk = fixed[ri] | fixed[col]
    A[msk] = np.where(ri[msk] == col[msk], 1.0, 0.0)
    dg = (ri == col) & ~fixed[col]
    A[dg] += 1e-10 * A[dg].max()
    L = N.lib()

    def sym(perm):
        po, par, cc = (np.zeros(n, np.int32) for _ in range(3))
        nz = C.c_longlong(0)
        rc = L.ifea_symbolic_analyze(n, N.ptr(N.i32(cp), C.c_int), N.ptr(N.i32(ri), C.c_int), N.ptr(N.i32(perm), C.c_int),
                                     None, N.ptr(po, C.c_int), N.ptr(par, C.c_int), N.ptr(cc, C.c_int), C.byref(nz),
                                     None)
        assert rc == 0, N.last_error()
        return po, par, cc, nz.value

    pd = O.Pardiso(n, cp, ri, want_perm=True)
    pd.factorize(A)
    mkl_perm = pd.perm()
    assert sorted(mkl_perm) == list(range(n))
    for perm in (mkl_perm, np.arange(n, dtype=np.int32)):
        po, par, cc, nzl = sym(perm)
        assert np.array_equal(po, perm)
        bpar, bcc = O.brute_symbolic(n, cp, ri, perm)
        assert np.array_equal(par, bpar) and np.array_equal(cc, bcc)
        assert nzl == int(bcc.sum())
        q = O.Pardiso(n, cp, ri, perm=perm)
        q.factorize(A)
        mkl_nnz = q.info()["nnz_l"]
        assert nzl <= mkl_nnz <= 1.05 * nzl, (nzl, mkl_nnz)
        # MKL's solution with the shared ordering equals the oracle Cholesky's with it
        b = np.random.default_rng(0).standard_normal(n)
        ch = O.OracleCholesky(n, cp, ri, perm)
        ch.factorize(A)
        x1, x2 = ch.solve(b), q.solve(b)
        assert np.linalg.norm(x1 - x2) <= 1e-7 * np.linalg.norm(x1)
    # the inverse of MKL's order is a different (and much worse) ordering: the seam would notice a transposed reading
    assert sym(np.argsort(mkl_perm).astype(np.int32))[3] > 2 * sym(mkl_perm)[3]


# ------------------------------------------------------------------ placement repair (TryUpdate :1455-1537)
def _grid(n):
    I, J, K = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    x = np.stack([I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()], 1)
    return x.astype(np.float64), kuhn_box(n, n, n).astype(np.int32)


def _quality(x, t):
    """6 sqrt(2) V / l2^1.5 with l2 the mean squared edge length: 1 for the regular tet, 0 when flat."""
    a, b, c, d = (x[t[:, k]] for k in range(4))
    e = [b - a, c - a, d - a, c - b, d - b, d - c]
    l2 = sum((v * v).sum(1) for v in e) / 6.0
    vol = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)) / 6.0
    return 6 * np.sqrt(2.0) * vol / l2 ** 1.5


@pytest.mark.parametrize("d", [(0.9, -0.9, 0.0), (0.0, 0.0, 0.95), (0.9, 0.9, 0.0), (0.3, 0.2, -0.97)])
def test_smooth_interior_returns_the_lone_interior_node_to_its_star_centroid(oracle_lib, d):
    """SmoothInterior (:1776-1838) on the 2x2x2 Kuhn box, whose only interior node is the centre: every other node is
    on the surface and stays put, and the centre's star is symmetric under the reflection through the centre, so the
    centroid of its tets' other vertices (counted with multiplicity) is the centre itself. Displacing the centre far
    enough to make tets bad (q < 0.05) therefore takes one full (lambda = 1) move straight back to it, and the next
    sweep finds nothing to do."""
    import oracle as O
    x, t = _grid(2)
    z = np.zeros(len(x), bool)
    c = int(np.nonzero(np.all(x == 1.0, axis=1))[0][0])
    xs = x.copy()
    xs[c] += np.array(d)
    assert (_quality(xs, t) < 0.05).any()
    o = O.Oracle(_cfg(2.0), x, t, z, z.copy(), 0.0)
    xo, moves = o.smooth_interior(xs, 0.05, 6)
    assert moves == 1
    assert np.abs(xo[c] - x[c]).max() <= 1e-15
    assert np.array_equal(np.delete(xo, c, 0), np.delete(xs, c, 0))
    # a bad tet made by a surface node alone is left to ImprovePlacement: nothing moves
    xs2 = x.copy()
    s = int(np.nonzero(np.all(x == (1.0, 1.0, 0.0), axis=1))[0][0])
    xs2[s, 2] = 0.97
    assert (_quality(xs2, t) < 0.05).any()
    xo2, moves2 = o.smooth_interior(xs2, 0.05, 6)
    assert moves2 == 1 and np.array_equal(xo2[s], xs2[s])      # only the centre (in a bad tet) moves
    assert np.abs(np.delete(xo2, [c, s], 0) - np.delete(xs2, [c, s], 0)).max() == 0.0


def placement_box_case():
    """A 4x4x4 unit Kuhn box between flat dies (bottom at z = 0, top closing at 10 mm/s from z = 4): the top node
    above the centre lags the die (the die overtakes it by 0.04 = 0.04 h in dt = 0.005) while the node below it rises
    to 0.04 under the die, flattening the tets between them. Returns (cfg, x, tets, contact, vg, gap, dt, n_top,
    n_below)."""
    n = 4
    x, t = _grid(n)
    g0 = 1.0
    cfg = StepConfig(size=4.0, h=1.0, bottom=DieSpec(0, 0, (0.0,) + (0.0,) * 7),
                     top=DieSpec(1, 0, (4.0 - g0,) + (0.0,) * 7))
    contact = np.where(x[:, 2] == 0, 1, np.where(x[:, 2] == n, 2, 0)).astype(np.int32)
    idx = lambda i, j, k: i + (n + 1) * (j + (n + 1) * k)
    top, below = idx(2, 2, 4), idx(2, 2, 3)
    dt = 0.005
    vg = np.zeros_like(x)
    vg[contact == 2, 2] = -cfg.ram
    vg[top, 2] = -0.2 * cfg.ram
    vg[below, 2] = 0.96 / dt
    return cfg, x, t, contact, vg, g0, dt, top, below


def test_improve_placement_flat_die_takes_the_best_candidate(oracle_lib):
    """ImprovePlacement (:1625-1679) on the flat-die case above. After placement the lagging top node sits on the die
    (z = 3.95) above its rising neighbour (z = 3.91), so its tets are bad. For a flat top die its candidates are known
    in closed form: the closest die point (its current place), the exit along the inward normal -z (the same point to
    the bisection's 0.1 h / 2^25), and, its penetration 0.04 being within 0.05 h, the partial pushes back from the
    predicted point z = 3.99 at f = 0, .25, .5, .75 (3.99, 3.98, 3.97, 3.96). It takes the candidate with the best
    worst-quality of its tets (computed here from the candidate list) and keeps its top contact (it still penetrates);
    SmoothInterior then moves the interior nodes of the remaining bad tets, and the step succeeds."""
    import oracle as O
    cfg, x, t, contact, vg, g0, dt, top, below = placement_box_case()
    z = np.zeros(len(x), bool)
    o = O.Oracle(cfg, x, t, z, z.copy(), g0)
    o.set_state(x=x, vg=vg, contact=contact, gap=g0)
    why, (reverts, smooth) = o.try_update(dt)
    assert why == "" and reverts == 1 and smooth >= 1
    s = o.state()
    assert s["gap"] == pytest.approx(g0 - cfg.ram * dt, abs=1e-15) and s["contact"][top] == 2
    # positions before the repair: the predicted ones, top contacts on the die at 3.95
    xp = x + dt * vg
    on_top = contact == 2
    xp[on_top, 2] = 3.95
    tets_n = np.nonzero(np.any(t == top, axis=1))[0]
    cands = [3.95, 3.99, 3.98, 3.97, 3.96]
    worst = []
    for zc in cands:
        xq = xp.copy()
        xq[top, 2] = zc
        worst.append(_quality(xq, t[tets_n]).min())
    best = int(np.argmax(worst))
    assert worst[best] > worst[0] + 1e-6
    assert s["x"][top, 2] == pytest.approx(cands[best], abs=1e-14)
    assert s["x"][top, 0] == 2.0 and s["x"][top, 1] == 2.0
    # the repair left no tet below the collapse threshold and no top node other than the lagging one off the die
    assert _quality(s["x"], t).min() > 0.05
    others = on_top.copy()
    others[top] = False
    assert np.all(s["x"][others, 2] == 3.95)
