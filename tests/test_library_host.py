"""Host-side checks of libifea_amd.so that need no GPU: every symbol of include/*.h is exported, the host symbolic
analysis (ordering, elimination tree, column counts, nnz(L)) equals the oracle's brute-force elimination bit for bit,
and the library fails loudly without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from impressionfea_amd import _native as N
from impressionfea_amd.mesh import TetMesh, build_pattern, kuhn_cylinder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(N.LIB_PATH):
        from impressionfea_amd import build
        build.build()
    return N.lib()


def _header_functions():
    names = set()
    for h in ("ifea_spd.h", "ifea_step.h", "ifea_mlp.h", "ifea_remesh.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(ifea_\w+)\s*\(", txt))
    return sorted(names)


def test_exports_every_declared_symbol(lib):
    decl = _header_functions()
    assert len(decl) >= 25
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing
    assert set(decl) == set(N.SPD_SYMBOLS) | set(N.STEP_SYMBOLS) | set(N.MLP_SYMBOLS) | set(N.REMESH_SYMBOLS)


def _sym(lib, n, cp, ri, perm=None, coords=None):
    po, par, cc = (np.zeros(n, np.int32) for _ in range(3))
    nz = C.c_longlong(0)
    st = np.zeros(6)
    pin = N.i32(perm) if perm is not None else None
    xyz = N.f64(coords) if coords is not None else None
    rc = lib.ifea_symbolic_analyze(n, N.ptr(N.i32(cp), C.c_int), N.ptr(N.i32(ri), C.c_int), N.ptr(pin, C.c_int),
                                   N.ptr(xyz, C.c_double), N.ptr(po, C.c_int), N.ptr(par, C.c_int),
                                   N.ptr(cc, C.c_int), C.byref(nz), N.ptr(st, C.c_double))
    assert rc == 0, N.last_error()
    return po, par, cc, nz.value, st


@pytest.mark.parametrize("nxy,nz,seed", [(3, 4, 0), (5, 6, 1), (10, 10, 0)])
def test_symbolic_bit_exact_vs_brute_force(lib, oracle_lib, nxy, nz, seed):
    import oracle as O
    x, t, ox, oy = kuhn_cylinder(40, 60, nxy, nz, seed=seed)
    m = TetMesh(x.shape[0], t, ox, oy)
    ns, nb, cp, ri = build_pattern(m)
    n = 3 * m.nn
    cases = [("nd", None, None), ("geo", None, np.repeat(x, 3, axis=0)),
             ("mindeg", O.min_degree_nodes(m.nn, ns, nb), None),
             ("random", np.random.default_rng(seed).permutation(n).astype(np.int32), None)]
    for name, perm, coords in cases:
        po, par, cc, nzl, st = _sym(lib, n, cp, ri, perm, coords)
        assert sorted(po) == list(range(n)), name
        if perm is not None:
            assert np.array_equal(po, perm), name  # a caller ordering is kept as given
        bpar, bcc = O.brute_symbolic(n, cp, ri, po)
        assert np.array_equal(par, bpar), name
        assert np.array_equal(cc, bcc), name
        assert nzl == int(bcc.sum()), name
        ch = O.OracleCholesky(n, cp, ri, po)
        assert ch.nnz == nzl, name


def test_symbolic_rejects_bad_input(lib):
    cp = np.array([0, 2, 4], np.int32)
    ri = np.array([0, 1, 0, 1], np.int32)
    po = np.zeros(2, np.int32)
    bad_perm = np.array([0, 0], np.int32)
    rc = lib.ifea_symbolic_analyze(2, N.ptr(cp, C.c_int), N.ptr(ri, C.c_int), N.ptr(bad_perm, C.c_int), None,
                                   N.ptr(po, C.c_int), None, None, None, None)
    assert rc == N.IFEA_ERR_PERM
    ri_bad = np.array([1, 0, 0, 1], np.int32)  # unsorted rows
    rc = lib.ifea_symbolic_analyze(2, N.ptr(cp, C.c_int), N.ptr(ri_bad, C.c_int), None, None, N.ptr(po, C.c_int),
                                   None, None, None, None)
    assert rc == N.IFEA_ERR_ARG


def test_no_device_fails_loudly(lib):
    """Without a HIP device the backend refuses to start (the factory's probe then falls back,
    SpdSolverFactory.cs:50-64); nothing silently computes on the CPU."""
    from impressionfea_amd import spd
    from impressionfea_amd.errors import EngineUnavailableError
    h = C.c_void_p()
    rc = lib.ifea_spd_create(C.byref(h))
    if rc == 0:
        lib.ifea_spd_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == N.IFEA_ERR_NO_DEVICE
    with pytest.raises(EngineUnavailableError):
        spd.HipSpdSolver()
    assert spd.self_test() is not None
