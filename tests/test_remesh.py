"""The remesh step on the host (CPU): the library's point locator and isosurface stuffing (include/ifea_remesh.h) and
the remesh pipeline of remesh.py, against the FP64 oracle's restatement of TetLocator, IsosurfaceStuffing and
RemeshCore (Forging3DSolver.cs:1958-2102) bit for bit, plus the properties the method guarantees (positive, manifold,
well-shaped tets; exact linear interpolation)."""
import os

import numpy as np
import pytest

from impressionfea_amd import remesh as RM
from impressionfea_amd.config import hexdie_production, upset_small
from impressionfea_amd.mesh import TetMesh, kuhn_cylinder
from impressionfea_amd.placement import tet_quality, tet_volumes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "remesh_states.npz")


@pytest.fixture(scope="module")
def lib():
    from impressionfea_amd import _native as N
    if not os.path.exists(N.LIB_PATH):
        from impressionfea_amd import build
        build.build()
    return N.lib()


def _quarter_cylinder():
    wl = upset_small(nxy=6, nz=8, quarter=True)
    x, t, ox, oy = kuhn_cylinder(wl.diameter, wl.height, wl.nxy, wl.nz, seed=3, quarter=True)
    return wl, x, t, ox, oy


def _hexdie_small():
    wl = hexdie_production(nxy=8, nz=12)
    x, t, ox, oy = wl.mesh()
    return wl, x, t, ox, oy


@pytest.mark.parametrize("case", ["quarter_cylinder", "hexdie_billet"])
def test_stuffing_bit_equal_to_oracle_and_well_shaped(lib, oracle_lib, case):
    """IsosurfaceStuffing.Stuff (IsosurfaceStuffing.cs:17-88, :115-294) of a billet at its element size: the library and
    the oracle give the same points and tets bit for bit, and the mesh has what the method guarantees: every tet
    positively oriented and above the 1e-6 h^3 sliver cut-off, a manifold (face-conforming) mesh, a quality floor
    (Labelle and Shewchuk bound the dihedral angles of this lattice and these alphas; the worst tet here is about 0.25)
    and the billet's volume to the faceting of a lattice of spacing h (a few per cent)."""
    import oracle as O
    wl, x, t, ox, oy = _quarter_cylinder() if case == "quarter_cylinder" else _hexdie_small()
    h, sym = wl.cfg.h, bool(ox.any())
    loc = RM.Locator(x, t, 2 * h)
    xs, ts = RM.stuff(loc, x.min(axis=0), x.max(axis=0), h, sym, sym)
    loc.close()
    px, pt = O.stuff(x, t, 2 * h, x.min(axis=0), x.max(axis=0), h, sym, sym)
    assert np.array_equal(xs, px) and np.array_equal(ts, pt)
    assert ts.shape[0] > 100
    v = tet_volumes(xs, ts)
    assert v.min() >= 1e-6 * h ** 3
    m = TetMesh(xs.shape[0], ts, (xs[:, 0] == 0) & sym, (xs[:, 1] == 0) & sym)    # raises when not manifold
    assert m.boundary_faces.shape[0] > 0 and np.all(np.bincount(ts.ravel(), minlength=xs.shape[0]) > 0)
    assert tet_quality(xs, ts).min() > 0.15
    v_old = tet_volumes(x, t).sum()
    assert abs(v.sum() - v_old) <= 0.08 * v_old
    if sym:
        # the symmetry planes are meshed flat: nodes exactly on them, and every node on the right side
        assert (xs[:, 0] == 0).any() and (xs[:, 1] == 0).any() and xs[:, :2].min() >= 0


def test_locator_bit_equal_to_oracle_and_exact_for_linear_fields(lib, oracle_lib):
    """TetLocator (TetMesh.cs:115-262): Find, its weights, FindNearest and Interpolate agree with the oracle bit for
    bit on points inside, near and outside the mesh. Closed forms: weights sum to 1 and reproduce the point; the
    interpolation of a linear field A p + b is exact inside (to rounding); FindNearest of an inside point is its
    containing tet; outside, Interpolate returns the field at the nearest node (brute force)."""
    import oracle as O
    wl, x, t, ox, oy = _quarter_cylinder()
    h = wl.cfg.h
    rng = np.random.default_rng(0)
    p = rng.uniform(x.min(axis=0) - 3, x.max(axis=0) + 3, (3000, 3))
    A, b = rng.standard_normal((3, 3)), rng.standard_normal(3)
    f = x @ A.T + b
    loc = RM.Locator(x, t, 2 * h)
    tet, w = loc.find(p)
    near = loc.find_nearest(p)
    ip = loc.interpolate(p, f)
    loc.close()
    ot, ow, on, oi = O.locate(x, t, 2 * h, p, f)
    assert np.array_equal(tet, ot) and np.array_equal(w, ow) and np.array_equal(near, on) and np.array_equal(ip, oi)
    inside = tet >= 0
    assert 0.2 < inside.mean() < 0.8
    assert np.abs(w[inside].sum(axis=1) - 1).max() < 1e-12
    corners = x[t[tet[inside]]]                                   # [k, 4, 3]
    assert np.abs(np.einsum("kj,kjc->kc", w[inside], corners) - p[inside]).max() < 1e-9
    scale = np.abs(f).max()
    assert np.abs(ip[inside] - (p[inside] @ A.T + b)).max() < 1e-11 * scale
    assert np.array_equal(near[inside], tet[inside])
    out = np.nonzero(~inside)[0][:50]
    for i in out:
        nearest = int(np.argmin(((x - p[i]) ** 2).sum(axis=1)))
        assert np.array_equal(ip[i], f[nearest])


def _state_case(name):
    """(cfg, mesh arrays, state dict, gap, vol0) for a remesh parity case."""
    if name == "barrel_quarter":
        wl, x, t, ox, oy = _quarter_cylinder()
        vol0 = float(np.cumsum(tet_volumes(x, t))[-1]) * wl.cfg.sym_factor
        H = wl.height
        s = 1 + 0.25 * np.sin(np.pi * x[:, 2] / H)
        xd = x.copy()
        xd[:, 0] *= s
        xd[:, 1] *= s
        xd[:, 2] *= 0.75
        contact = np.where(x[:, 2] == 0, 1, np.where(x[:, 2] == H, 2, 0)).astype(np.int32)
        rng = np.random.default_rng(1)
        ne, nn = t.shape[0], x.shape[0]
        st = dict(eps=rng.random(ne), dmg=np.zeros(ne), sig=50 + rng.random(ne), rate=rng.random(ne),
                  mean=rng.standard_normal(ne), vg=rng.standard_normal((nn, 3)), contact=contact)
        return wl.cfg, (xd, t, ox, oy), st, 0.75 * H - 20.0, vol0
    z = np.load(GOLDEN, allow_pickle=False)
    k = name[1:]
    pre = f"s{k}_"
    wl = hexdie_production(nxy=8, nz=12)
    st = dict(eps=z[pre + "eps"], dmg=z[pre + "dmg"], sig=z[pre + "sig"], rate=z[pre + "rate"], mean=z[pre + "mean"],
              vg=z[pre + "vg"], contact=z[pre + "contact"])
    return wl.cfg, (z[pre + "x"], z[pre + "tets"], z[pre + "on_x0"], z[pre + "on_y0"]), st, float(z[pre + "gap"]), \
        float(z["vol0"])


@pytest.mark.parametrize("name", ["barrel_quarter", "s26", "s49"])
def test_remesh_core_bit_equal_to_oracle(lib, oracle_lib, name):
    """RemeshCore (:1958-2102): stuffing, contact from the dies, volume compensation, ImprovePlacement at q 0.1, the
    revert sweeps, SmoothInterior / UntangleSurface / SmoothInterior, the acceptance rule and the remap of element
    history (nearest old tet of each new centroid) and velocities (interpolated). On a barrel-shaped quarter upset and
    on the oracle's own closed-die states at the two remeshes of its run (flash forming; fixture written by
    tests/golden/make_remesh_states.py), the host pipeline and the oracle give the same mesh, positions, contact sides
    and fields bit for bit, and the same counts of repair moves."""
    import oracle as O
    cfg, (x, t, ox, oy), st, gap, vol0 = _state_case(name)
    o = O.Oracle(cfg, x, t, ox, oy, gap)
    o.set_state(x=x, vg=st["vg"], contact=st["contact"], eps=st["eps"], elSig=st["sig"], gap=gap)
    o.set_element_state(elRate=st["rate"], elMean=st["mean"])
    symX, symY = bool(np.any(ox)), bool(np.any(oy))
    o.set_remesh_info(vol0, symX, symY)
    if not np.any(st["dmg"]):
        dmg = np.zeros(t.shape[0])
    else:
        dmg = st["dmg"]
        o.set_damage(dmg)
    m = TetMesh(x.shape[0], t, np.asarray(ox, bool), np.asarray(oy, bool))
    xp = np.array(x, float)
    if name != "barrel_quarter":
        # as at the start of the increment that remeshed: the trigger's interior smoothing first, on both sides
        why = RM.remesh_trigger(m, xp, 0.12, cfg.h)
        code = o.remesh_trigger(0.12)
        assert (why is None) == (code == 0)
        assert np.array_equal(xp, o.state()["x"])
    state = dict(eps=st["eps"], dmg=dmg, sig=st["sig"], rate=st["rate"], mean=st["mean"], vg=st["vg"])
    res = RM.remesh_core(cfg, m, xp, state, gap, vol0, cfg.sym_factor, cfg.constants()["sigRef"], symX, symY, 0)
    ok, ost = o.remesh(0)
    assert ok == (res is not None)
    assert res is not None
    for k in ("improve", "reverts", "smooth", "untangle"):
        assert res.stats[k] == ost[k], k
    assert res.stats["q_new"] == ost["q_new"] and res.stats["q_old"] == ost["q_old"]
    s1, es = o.state(), o.element_state()
    tt, a, b = o.mesh()
    assert np.array_equal(res.tets, tt) and np.array_equal(res.onX0, a) and np.array_equal(res.onY0, b)
    assert np.array_equal(res.x, s1["x"])
    assert np.array_equal(res.contact, s1["contact"])
    assert np.array_equal(res.vg, s1["vg"])
    assert np.array_equal(res.eps, s1["eps"]) and np.array_equal(res.sig, s1["elSig"])
    assert np.array_equal(res.rate, es["elRate"]) and np.array_equal(res.mean, es["elMean"])
    assert np.array_equal(res.dmg, es["dmg"])
    assert res.stats["q_new"] > 0.02
    if name == "s49":
        assert res.stats["improve"] + res.stats["reverts"] + res.stats["untangle"] > 0   # thin flash: repairs act


def test_remesh_trigger_matches_oracle(lib, oracle_lib):
    """RemeshTrigger (:1867-1880) on the closed-die billet's initial Kuhn mesh (worst tet 0.093 < 0.12): the interior
    smoothing moves the same nodes to the same places as the oracle's, and both then ask for a remesh on quality."""
    import oracle as O
    wl, x, t, ox, oy = _hexdie_small()
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0)
    o.initial(float(x[:, 2].min()))
    xs = o.state()["x"]
    m = TetMesh(x.shape[0], t, ox, oy)
    xp = xs.copy()
    why = RM.remesh_trigger(m, xp, 0.12, wl.cfg.h)
    assert o.remesh_trigger(0.12) == 1 and why.startswith("element quality")
    assert np.array_equal(xp, o.state()["x"])
    # a fresh, regular box mesh needs nothing
    from impressionfea_amd.mesh import kuhn_box
    n = 3
    I, J, K = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    xb = np.stack([I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()], 1) * 2.0
    tb = kuhn_box(n, n, n).astype(np.int32)
    z = np.zeros(xb.shape[0], bool)
    assert RM.remesh_trigger(TetMesh(xb.shape[0], tb, z, z), xb.astype(float), 0.12, 2.0) is None


def test_untangle_surface_and_volume_compensation_match_oracle(lib, oracle_lib):
    """UntangleSurface (:1698-1772) and CompensateVolume (:2108-2145) on a quarter billet whose top-surface nodes were
    pushed down into the layer below (flattened, inverted tets at the top, as at closure): both move the same nodes to
    the same places as the oracle's restatement; untangling acts (the stuffed, i.e. original, positions restore the
    tets), releases the nodes it moves and lifts the worst tet above 0.03; compensation restores the lost volume to
    within the one-h cap."""
    import oracle as O
    from impressionfea_amd import placement as PL
    wl, x, t, ox, oy = _quarter_cylinder()
    cfg = wl.cfg
    H = wl.height
    o = O.Oracle(cfg, x, t, ox, oy, wl.gap0)
    vol0 = float(np.cumsum(tet_volumes(x, t))[-1]) * cfg.sym_factor
    o.set_remesh_info(vol0, True, True)
    m = TetMesh(x.shape[0], t, ox, oy)
    top = np.nonzero(x[:, 2] == H)[0]
    xd = x.copy()
    xd[top[::2], 2] -= 1.1 * H / wl.nz            # every other top node below the next layer
    contact = np.zeros(x.shape[0], np.int32)
    contact[top] = 2
    assert tet_quality(xd, t).min() < 0.03
    xp, cp = xd.copy(), contact.copy()
    n = PL.untangle_surface(m, xp, cp, x, 0.03, 6, cfg, wl.gap0, PL.host_bad_tets(m))
    xo, co, no = o.untangle_surface(xd, contact, x, 0.03, 6)
    assert n == no and n > 0
    assert np.array_equal(xp, xo) and np.array_equal(cp, co)
    assert tet_quality(xp, t).min() > 0.03
    assert np.all(cp[np.any(xp != xd, axis=1)] == 0)
    # volume compensation on the shrunk billet (free surface only: contact nodes stay)
    xs = x.copy()
    r = np.hypot(xs[:, 0], xs[:, 1])
    xs[:, :2] *= (1 - 0.02 * (r > 0))[:, None]
    cz = np.where(x[:, 2] == 0, 1, np.where(x[:, 2] == H, 2, 0)).astype(np.int32)
    xq = xs.copy()
    PL.compensate_volume(m, xq, cz, vol0, cfg.sym_factor, cfg.h, cfg, wl.gap0)
    xo2 = o.compensate_volume(xs, cz)
    assert np.array_equal(xq, xo2)
    v_before, v_after = tet_volumes(xs, t).sum(), tet_volumes(xq, t).sum()
    assert abs(v_after - vol0 / 4) < 0.2 * abs(v_before - vol0 / 4)
    assert np.array_equal(xq[cz != 0], xs[cz != 0])
