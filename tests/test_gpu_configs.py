"""The production workloads on the GPU: the 360,555-DOF closed-die hex upset against the oracle's committed loads,
the 973,299-DOF connecting-rod die at full size through properties that need no CPU factorization, and a reduced
held-out evaluation of the learned warm start.

Each test prints what it measured (loads, errors, iteration counts) so a run records the numbers, not only a pass.
"""
import ctypes as C
import json
import os
import time

import numpy as np
import pytest
import scipy.sparse as sp

from impressionfea_amd import _native as N
from impressionfea_amd import spd
from impressionfea_amd.config import conrod_production, hexdie_production
from impressionfea_amd.forging3d import Forging3DSolver

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_hexdie_two_increments_against_oracle_loads():
    """The first two increments of the hex-die workload (120,185 nodes, 677,376 tets, 360,555 DOF), plain FP64, mixed
    precision, and the benchmark's mixed method (mixed precision, learned start judged against the reference's
    starts, adaptive first solve, PCG stopped at 1e-2), against the ram loads of the FP64 CPU oracle with MKL PARDISO
    committed in tests/golden/hexdie_loads.json (written by tests/golden/make_hexdie_loads.py on the same seeded
    mesh): FP64 within 1e-6, the mixed ones within 1e-3 (the 0.1 % bar)."""
    from impressionfea_amd import warmstart
    gold = json.load(open(os.path.join(HERE, "golden", "hexdie_loads.json")))
    wl = hexdie_production()
    x, t, _, _ = wl.mesh()
    assert x.shape[0] == gold["nodes"] and t.shape[0] == gold["tets"]
    assert float(np.abs(x).sum()) == pytest.approx(gold["x_checksum"], rel=1e-15)
    ref = [r["load"] for r in gold["increments"]]
    bench_mixed = dict(warm_start=warmstart.load_default(), first_solve_policy=True, pcg_tol=1e-2)
    for precision, bar, kw in ((0, 1e-6, {}), (1, 1e-3, {}), (1, 1e-3, bench_mixed)):
        t0 = time.time()
        s = Forging3DSolver.from_workload(wl, precision=precision, **kw)
        for k in range(2):
            r = s.step()
            err = abs(r.load / ref[k] - 1)
            print(f"precision {precision}{' (benchmark mixed method)' if kw else ''} increment {k}: load "
                  f"{r.load:.9e} N, oracle {ref[k]:.9e} N, rel err "
                  f"{err:.2e}, Newton {r.iterations}, factorizations {r.stats['factorizations']}, "
                  f"fallbacks {r.stats['fp64_fallbacks']} ({time.time() - t0:.1f} s)", flush=True)
            assert r.converged
            assert err <= bar
        s.close()


def _device_used():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value, total.value


def test_conrod_full_size_properties():
    """The connecting-rod die at full size (324,433 nodes, 1,843,200 tets, 973,299 DOF): the first Picard system as
    SolveVelocity forms it, assembled on the GPU, factored and solved through the ISpdSolver slot. Properties that need
    no CPU factorization of this size: the normwise backward error |A x - b| / (|A| |x| + |b|) (scipy's SpMV on the
    host) of a backward-stable FP64 factorization, the slot's nnz(L) and column counts equal to the host-only symbolic
    analysis of the same ordering, and the device memory of the process below the MI355X's 288 GB."""
    t0 = time.time()
    wl = conrod_production()
    s = Forging3DSolver.from_workload(wl)
    s.prepare()
    val, b = s.linear_system(newton=False)
    n = s.ndof
    print(f"conrod: {s.nn} nodes, {s.ne} tets, {n} DOF, {s.nnz} entries; system assembled ({time.time() - t0:.1f} s)",
          flush=True)
    with spd.HipSpdSolver() as h:
        h.set_coordinates(np.repeat(s.x, 3, axis=0))
        h.Analyze(n, s.colPtr, s.rowIdx)
        perm, parent, cc = h.symbolic()
        nnzl = h.FactorNonZeros
        print(f"  analyzed: nnz(L) {nnzl:.4e} ({time.time() - t0:.1f} s)", flush=True)
        h.Factorize(val)
        x = h.Solve(b)
        used, total = _device_used()
        print(f"  factored and solved ({time.time() - t0:.1f} s); device memory {used / 2**30:.1f} GiB of "
              f"{total / 2**30:.1f} GiB", flush=True)
    A = sp.csc_matrix((val, s.rowIdx, s.colPtr), shape=(n, n))
    res = np.linalg.norm(A @ x - b)
    anorm = abs(A).sum(axis=0).max()        # ||A||_1
    eta = res / (anorm * np.linalg.norm(x) + np.linalg.norm(b))
    print(f"  relative residual {res / np.linalg.norm(b):.2e}, normwise backward error {eta:.2e}", flush=True)
    # a backward-stable FP64 factorization gives a normwise backward error of a small multiple of n eps |A|; the
    # relative residual |r| / |b| is larger by the condition of the penalty system (~1e10 at this penalty)
    assert np.isfinite(x).all() and eta <= 1e-14 and res <= 1e-6 * np.linalg.norm(b)
    # the host-only analysis of the slot's ordering gives the same structure
    po, par2, cc2 = (np.zeros(n, np.int32) for _ in range(3))
    nz = C.c_longlong(0)
    rc = N.lib().ifea_symbolic_analyze(n, N.ptr(N.i32(s.colPtr), C.c_int), N.ptr(N.i32(s.rowIdx), C.c_int),
                                       N.ptr(N.i32(perm), C.c_int), None, N.ptr(po, C.c_int), N.ptr(par2, C.c_int),
                                       N.ptr(cc2, C.c_int), C.byref(nz), None)
    assert rc == 0, N.last_error()
    assert nz.value == nnzl and np.array_equal(cc2, cc) and np.array_equal(par2, parent)
    assert used < 288e9 <= total + 1e9
    s.close()


def test_learned_warm_start_reduced_held_out():
    """A reduced config-4 evaluation on the GPU: the network trained by the project's own HIP kernels on four seeded
    upset billets (10 increments each), then one held-out billet stepped from the previous-increment start (the
    reference's default), the trend start (its ExtrapolateStart), the energy judge without the network, and the
    learned start under the judge. Training must lower the loss below the plain trend start's; the judged starts must
    not cost Newton iterations against the previous start nor move the ram load beyond the 0.1 % bar. The printed
    table is what the network gains over the method without it."""
    from impressionfea_amd import train_warmstart as T
    from impressionfea_amd.mlp import LearnedWarmStart
    import tempfile
    out = os.path.join(tempfile.mkdtemp(), "ws.npz")
    t0 = time.time()
    logs = []
    hist, base = T.train(seeds=[0, 1, 2, 3], epochs=20, increments=10, out=out, log=logs.append)
    print(f"trained in {time.time() - t0:.1f} s: loss {hist[0]:.3e} -> {hist[-1]:.3e}, trend start {base:.3e}")
    assert hist[-1] < base
    rows = T.evaluate(seeds=[100], weights=out, increments=10, log=logs.append)
    r = rows[0]
    print(json.dumps(r))
    assert r["iterations_judged_without_network"] <= r["iterations_previous"]
    assert r["iterations_learned"] <= r["iterations_previous"]
    assert r["max_load_rel_diff_learned"] <= 1e-3 and r["max_load_rel_diff_trend"] <= 1e-3
    assert LearnedWarmStart.load(out) is not None
