"""Benchmark: stroke increments per second of the 3D solve step at ~360k DOF on one MI355X.

Workload (default): the closed-die hex-head upset, billet D 30 x H 45 mm into a hexagonal cavity (across-flats
38 mm, 2 mm flash land), 42x42x64 Kuhn cells = 120,185 nodes / 677,376 tets / 360,555 DOF, synthetic seeded mesh.
Four arms run from the same initial state:
  fp64            plain FP64 supernodal factor, previous-increment start, PCG on the cached factor at every solve: the
                  reference's algorithm (Forging3DSolver.cs:854-1019) on the GPU
  fp64_trend      the same with the trend start 2 v_n - v_{n-1} (the reference's ExtrapolateStart option)
  mixed_previous  mixed precision (FP32 MFMA factor, or at a penalty beyond FP32 the FP64 factor with an FP32 copy for
                  the triangular sweeps; FP64 residual and PCG refinement) with the previous-increment start
  mixed           the full method: mixed precision, the learned start judged by the solver's energy against the
                  previous and trend starts, and the adaptive first solve (an increment's first solve refactors
                  directly once PCG on the previous increment's factor has missed its budget twice in a row)
A "step" is one stroke increment: PrepareStep (frames on the host, geometry/ANP/friction on the GPU) + SolveVelocity +
FinalEvaluation + strain update, followed by the host geometry update (TryUpdate), which is identical in every arm.
`value` is the mixed arm's increments per second of wall clock over the timed steps (TryUpdate included; the rate
without it is reported too); the other arms, every arm's load error against fp64, the precision each arm actually
computed in, the roofline of the dominant kernel (the outer MFMA supernode update, HIP events around every launch on
the library's stream) and a host CPU baseline (the FP64 oracle with MKL PARDISO on the host cores, timed piecewise on a
bounded sample and scaled by the fp64 arm's per-increment counts: a model, labelled as such) are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense MFMA TFLOP/s
HBM_PEAK = 8000.0                      # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# The arms: (factor precision, start, adaptive first-solve policy, PCG tolerance). "fp64" is the reference's algorithm
# on the GPU (PCG to 1e-7, Forging3DSolver.cs:984); the two middle arms separate the effect of the start from that of
# the precision; "mixed" is the full method, whose Newton iteration is inexact: PCG stops at a relative residual of
# 1e-2 (Newton's convergence test on the velocity step is the reference's).
ARMS = {
    "fp64": dict(precision=0, start="previous", first_solve=False, pcg_tol=1e-7),
    "fp64_trend": dict(precision=0, start="trend", first_solve=False, pcg_tol=1e-7),
    "mixed_previous": dict(precision=1, start="previous", first_solve=False, pcg_tol=1e-7),
    "mixed": dict(precision=1, start="learned", first_solve=True, pcg_tol=1e-2),
}


def run_arm(wl, precision, steps, warmup, warm_start=None, profile=True, start="previous", first_solve=False,
            pcg_tol=1e-7):
    from dataclasses import replace
    from impressionfea_amd.forging3d import Forging3DSolver
    if start == "trend":
        wl = replace(wl, cfg=replace(wl.cfg, extrapolate_start=True))
    s = Forging3DSolver.from_workload(wl, precision=precision, warm_start=warm_start if start == "learned" else None,
                                      first_solve_policy=first_solve, pcg_tol=pcg_tol)
    s.set_params(profile=profile)
    loads, ms_solve, ms_update, stats, all_loads = [], [], [], [], []
    for k in range(warmup):
        r = s.step()
        all_loads.append(r.load)
        log(f"  warmup {k}: load {r.load:.6e} N it {r.iterations} fact {r.stats['factorizations']} "
            f"{r.ms_solve:.0f} ms")
    s.info(reset=True)
    t0 = time.perf_counter()
    for k in range(steps):
        r = s.step()
        loads.append(r.load)
        all_loads.append(r.load)
        ms_solve.append(r.ms_solve)
        ms_update.append(r.ms_update)
        stats.append(r.stats)
        log(f"  step {k}: load {r.load:.6e} N it {r.iterations} fact {r.stats['factorizations']} "
            f"pcg {r.stats['pcg_iterations']} solves {r.stats['solves']} fb {r.stats['fp64_fallbacks']} "
            f"skip {r.stats['first_solve_skips']} start {r.stats.get('start')} {r.ms_solve:.0f} ms "
            f"(+{r.ms_update:.0f} ms update, repair moves {r.stats['repair'][0]}/{r.stats['repair'][1]})")
    wall = time.perf_counter() - t0  # every step returns after its device work (results are read back)
    info = s.info(reset=True)
    used = device_memory_used()
    s.close()
    return dict(loads=loads, ms_solve=ms_solve, ms_update=ms_update, stats=stats, wall=wall, info=info,
                device_used=used, all_loads=all_loads)


def device_memory_used():
    """Bytes in use on the current device (total - free, hipMemGetInfo), or None when it cannot be read."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
        free, total = C.c_size_t(), C.c_size_t()
        if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
            return None
        return dict(used=total.value - free.value, total=total.value)
    except OSError:
        return None


def cpu_baseline(wl, gpu_stats, threads, nnz_gpu=0):
    """FP64 oracle (C++ restatement, OpenMP assembly) with MKL PARDISO on the host cores, on a bounded sample of the
    same workload: one PrepareStep, one tangent assembly, one residual assembly, one PARDISO factorisation and one
    solve; scaled to increments/s with the fp64 arm's per-increment counts of assemblies, factorisations and solves."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    if not O.pardiso_available():
        return None
    O.set_threads(threads)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    log(f"cpu baseline: oracle + MKL PARDISO on {threads} host threads")
    x, t, ox, oy = wl.mesh()
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0, linear=1)
    o.initial(float(x[:, 2].min()))
    t0 = time.perf_counter()
    o.prepare()
    t_prep = time.perf_counter() - t0
    v = o.v()
    t0 = time.perf_counter()
    e, p, r, val = o.assemble(v, True, True)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.assemble(v, False, False)
    t_res = time.perf_counter() - t0
    fr = o.frames()
    ns, nb, cp, ri = o.pattern()
    n = o.ndof
    fixed = fr["fixed"]
    col = np.repeat(np.arange(n), np.diff(cp))
    msk = fixed[ri] | fixed[col]
    val[msk] = np.where(ri[msk] == col[msk], 1.0, 0.0)
    dg = (ri == col) & ~fixed[col]
    val[dg] += 1e-10 * val[dg].max()
    b = -np.where(fixed, 0.0, r)
    log(f"  PrepareStep {t_prep:.2f}s, tangent assembly {t_asm:.2f}s, residual assembly {t_res:.2f}s")
    t0 = time.perf_counter()
    pd = O.Pardiso(n, cp, ri)
    t_an = time.perf_counter() - t0
    log(f"  PARDISO analyse {t_an:.1f}s")
    t0 = time.perf_counter()
    pd.factorize(val)
    t_fac = time.perf_counter() - t0
    log(f"  PARDISO factor {t_fac:.1f}s")
    t0 = time.perf_counter()
    pd.solve(b)
    t_sol = time.perf_counter() - t0
    info = pd.info()
    # MKL (LP64) reports nnz(L) in a 32-bit field: past 2^32 it is the count modulo 2^32
    mkl_counts = f"{info['nnz_l']:.3e} nnz(L)" + (" modulo 2^32, MKL's 32-bit field" if nnz_gpu >= 2 ** 32 else "")
    k = len(gpu_stats)
    asm_m = sum(s["iterations"] for s in gpu_stats) / k
    asm_r = sum(s["assemblies"] for s in gpu_stats) / k - asm_m + 1  # line-search + final residual assemblies
    fac = sum(s["factorizations"] for s in gpu_stats) / k
    sol = sum(s["solves"] for s in gpu_stats) / k
    spmv = sum(s["pcg_iterations"] for s in gpu_stats) / k
    per_inc = t_prep + asm_m * t_asm + asm_r * t_res + fac * t_fac + sol * t_sol + spmv * 0.25 * t_res
    model = ""
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return dict(value=1.0 / per_inc, unit="increments/s", cores=threads, kind="port", cpu=model,
                method=("modelled: the host's piecewise times of one increment's parts, multiplied by the fp64 arm's "
                        "per-increment counts (an end-to-end oracle increment with MKL PARDISO takes longer: "
                        "tests/golden/make_hexdie_loads.py records whole increments)"),
                sample=(f"{wl.name} {3 * x.shape[0]:,} DOF, one increment's work measured piecewise on the host: PrepareStep "
                        f"{t_prep:.2f}s, tangent assembly {t_asm:.2f}s, residual assembly {t_res:.2f}s, MKL PARDISO "
                        f"analyse {t_an:.1f}s (not per increment) / factor {t_fac:.2f}s ({mkl_counts}) / solve "
                        f"{t_sol:.2f}s; scaled by the fp64 arm's per-increment "
                        f"counts (assemblies {asm_m:.1f}+{asm_r:.1f}, factorisations {fac:.2f}, solves {sol:.1f})"))


def oracle_parity(wl, res, n, threads):
    """The FP64 oracle (C++ restatement of the reference solver, MKL PARDISO with the reference's settings) on the
    first n increments of the same workload, from the same initial state; the ram load of every GPU arm's matching
    increment (warmup included) is compared with it. Test infrastructure used as the checker: bounded by n."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    if not O.pardiso_available():
        log("oracle parity: MKL PARDISO not available, skipped")
        return None
    O.set_threads(threads)
    x, t, ox, oy = wl.mesh()
    o = O.Oracle(wl.cfg, x, t, ox, oy, wl.gap0, linear=1)
    o.initial(float(x[:, 2].min()))
    loads, its, secs = [], [], []
    for k in range(n):
        t0 = time.perf_counter()
        r = o.step()
        secs.append(time.perf_counter() - t0)
        loads.append(r["load"])
        its.append(r["iterations"])
        log(f"  oracle increment {k}: load {r['load']:.9e} N it {r['iterations']} fact {r['factorizations']} "
            f"{secs[-1]:.1f} s")
    out = dict(increments=n, oracle="FP64 C++ restatement + MKL PARDISO (iparm as the reference)", threads=threads,
               oracle_loads=loads, oracle_iterations=its, oracle_seconds=[round(v, 1) for v in secs])
    for arm, v in res.items():
        g = v["all_loads"][:n]
        out[f"{arm}_loads"] = g
        out[f"{arm}_rel_err"] = [abs(a - b) / abs(b) for a, b in zip(g, loads)]
    return out


def copy_probe(nbytes: int, reps: int = 3):
    """Device-to-device copies of a known size (the calibration kernel for HBM byte counters)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8,
                    help="timed stroke increments per arm (after the warmup increments); the first increments after "
                         "the cold start, with the contact set still developing, are the slowest")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hexdie", choices=["hexdie", "upset", "conrod"],
                    help="hexdie: the 360k-DOF hex-head closed die (the headline line); conrod: the ~1M-DOF "
                         "connecting-rod closed die (peak device memory against the 288 GB of one MI355X); upset: "
                         "the small upset")
    ap.add_argument("--arms", default="all",
                    help="comma-separated arms out of " + ", ".join(ARMS) + "; 'all' (default) runs every arm, "
                         "'both' only fp64 and mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the triangular sweeps one by one instead of replaying them as a HIP graph "
                         "(for kernel-trace profilers)")
    ap.add_argument("--oracle-increments", type=int, default=0, metavar="N",
                    help="also run the FP64 CPU oracle (MKL PARDISO) on the first N increments and report every "
                         "arm's ram-load error against it (minutes per increment at 360k DOF)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of this size (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.no_graph:
        os.environ["IFEA_SWEEP_GRAPH"] = "0"
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    from impressionfea_amd.config import conrod_production, hexdie_production, upset_small
    from impressionfea_amd import warmstart
    wl = {"hexdie": hexdie_production, "conrod": conrod_production, "upset": upset_small}[args.workload]()
    ws = warmstart.load_default()
    log(f"workload {wl.name}: {wl.nxy}x{wl.nxy}x{wl.nz} cells; learned warm start: {'yes' if ws else 'no'}")
    arms = list(ARMS) if args.arms == "all" else (["fp64", "mixed"] if args.arms == "both" else args.arms.split(","))
    for a in arms:
        if a not in ARMS:
            sys.exit(f"unknown arm {a!r}; choose from {', '.join(ARMS)}")
    res = {}
    for a in arms:
        spec = ARMS[a]
        log(f"arm {a}: precision {spec['precision']}, {spec['start']} start, adaptive first solve "
            f"{'on' if spec['first_solve'] else 'off'}, PCG tolerance {spec['pcg_tol']:g}")
        res[a] = run_arm(wl, spec["precision"], args.steps, args.warmup, warm_start=ws, start=spec["start"],
                         first_solve=spec["first_solve"], pcg_tol=spec["pcg_tol"])
    head = res.get("mixed") or res[arms[-1]]
    ips = {k: len(v["ms_solve"]) / v["wall"] for k, v in res.items()}                  # whole increments, wall clock
    ips_solve = {k: len(v["ms_solve"]) / (sum(v["ms_solve"]) / 1e3) for k, v in res.items()}  # without TryUpdate
    value = ips["mixed"] if "mixed" in ips else ips[arms[-1]]
    ms_step = 1e3 / value
    load_err = {}
    if "fp64" in res:
        a = np.array(res["fp64"]["loads"])
        for k, v in res.items():
            if k != "fp64":
                load_err[k] = float(np.max(np.abs(np.array(v["loads"]) - a) / np.abs(a)))
    # the arithmetic the timed arms actually used: factorizations and triangular sweeps by precision
    precision_used = {k: dict(fp32_factorizations=v["info"]["fp32_factorizations"],
                              fp64_factorizations=v["info"]["factorizations"] - v["info"]["fp32_factorizations"],
                              fp32_sweeps=v["info"]["fp32_sweeps"],
                              fp64_sweeps=v["info"]["sweeps"] - v["info"]["fp32_sweeps"]) for k, v in res.items()}
    hp = precision_used["mixed" if "mixed" in res else arms[-1]]
    dtype = ("fp32 factor + fp64" if hp["fp32_factorizations"] else "fp64") + (
        " (fp32 sweeps)" if hp["fp32_sweeps"] else "")
    # roofline of the dominant kernel: the outer supernode update k_syrk_big (MFMA), from HIP events around every
    # launch on the library's stream; split by the precision the factor was actually computed in
    info = head["info"]
    parts = {"fp32": (info["outer_launches_fp32"], info["outer_ms_fp32"], info["outer_flops_fp32"]),
             "fp64": (info["outer_launches"] - info["outer_launches_fp32"], info["outer_ms"] - info["outer_ms_fp32"],
                      info["outer_flops"] - info["outer_flops_fp32"])}
    roofs = {}
    for prec, (nl, ms, fl) in parts.items():
        if nl <= 0 or ms <= 0:
            continue
        ach = fl / (ms / 1e3) / 1e12
        roofs[prec] = dict(bound="mfma", achieved=round(ach, 3), peak=PEAK[prec], unit="TFLOP/s",
                           frac=round(ach / PEAK[prec], 4), traffic=None,
                           kernel=f"k_syrk_big<{'float' if prec == 'fp32' else 'double'}>", launches=nl,
                           avg_launch_us=round(1e3 * ms / nl, 2), flops_per_launch=round(fl / nl),
                           total_ms=round(ms, 1))
    # HBM bytes per launch of the same kernel from the two PMC passes (FETCH_SIZE, WRITE_SIZE) summarised in profiles/
    traffic = {}
    pmc = os.path.join(ROOT, "profiles", "k_syrk_big_hexdie_pmc.md")
    if args.workload == "hexdie" and os.path.exists(pmc):
        for line in open(pmc):
            for prec in ("fp64", "fp32"):
                if line.startswith(f"traffic_bytes_per_launch_{prec}:"):
                    traffic[prec] = float(line.split(":", 1)[1])
    for prec, r in roofs.items():
        r["traffic"] = traffic.get(prec)
    roof = None
    if roofs:
        dom = max(roofs, key=lambda k: roofs[k]["total_ms"])
        roof = dict(roofs[dom])
        other = [v for k, v in roofs.items() if k != dom]
        if other:
            roof["other_precision"] = other[0]
        if info["syrk_launches"]:
            roof["inner_updates"] = dict(kernel="k_syrk", launches=info["syrk_launches"],
                                         total_ms=round(info["syrk_ms"], 1),
                                         achieved=round(info["syrk_flops"] / (info["syrk_ms"] / 1e3) / 1e12, 3))
    sweep = None
    if info["sweep_ms"] > 0:
        bw = info["sweep_bytes"] / (info["sweep_ms"] / 1e3) / 1e9
        sweep = dict(bound="hbm", achieved=round(bw, 1), peak=HBM_PEAK, unit="GB/s", frac=round(bw / HBM_PEAK, 4),
                     solves=info["sweeps"], fp32_solves=info["fp32_sweeps"],
                     avg_ms=round(info["sweep_ms"] / max(1, info["sweeps"]), 3))
    ndof = 3 * ((wl.nxy + 1) ** 2 * (wl.nz + 1))
    memory = {k: dict(factor_bytes=v["info"]["device_bytes"],
                      process_bytes=v["device_used"]["used"] if v["device_used"] else None,
                      device_total=v["device_used"]["total"] if v["device_used"] else None)
              for k, v in res.items()}
    if args.workload == "hexdie":
        metric = "stroke increments/s @360k DOF: plain FP64 vs mixed+warm-start, <=0.1% load err"
        desc = f"{wl.name} closed-die hex upset, {wl.nxy}x{wl.nxy}x{wl.nz} Kuhn cells, 360,555 DOF, FP64 residual"
    elif args.workload == "conrod":
        metric = "stroke increments/s @1M DOF con-rod: plain FP64 vs mixed, peak factor memory vs 288 GB"
        desc = (f"{wl.name} connecting-rod closed die, billet lying along x, {wl.nxy}x{wl.nxy}x{wl.nz} Kuhn cells, "
                f"{ndof:,} DOF, FP64 residual")
    else:
        metric, desc = "stroke increments/s (small upset)", wl.name
    out = {
        "metric": metric,
        "value": round(value, 4), "unit": "increments/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic",
        "config": {"workload": desc, "dof": ndof,
                   "arms": {k: round(v, 4) for k, v in ips.items()},
                   "arm_definitions": {k: ARMS[k] for k in res},
                   "arms_without_host_update": {k: round(v, 4) for k, v in ips_solve.items()},
                   "speedup_vs_fp64": ({k: round(v / ips["fp64"], 3) for k, v in ips.items() if k != "fp64"}
                                       if "fp64" in ips else None),
                   "newton_iterations": {k: [st["iterations"] for st in v["stats"]] for k, v in res.items()},
                   "per_increment_means": {k: {q: round(float(np.mean([st[q] for st in v["stats"]])), 2)
                                               for q in ("factorizations", "solves", "pcg_iterations",
                                                         "first_solve_skips", "fp64_fallbacks")}
                                           for k, v in res.items()},
                   "starts": {k: {nm: sum(st.get("start") == nm for st in v["stats"])
                                  for nm in ("previous", "trend", "learned")} for k, v in res.items()},
                   "precision_used": precision_used,
                   "load_rel_err_vs_fp64": load_err,
                   "learned_warm_start": bool(ws),
                   "nnz_L": info["nnz_l"], "factor_flops": info["flops"],
                   "ms_update_host": round(float(np.mean(head["ms_update"])), 1),
                   "device_memory": memory},
        "roofline": roof,
        "triangular_sweeps": sweep,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and "fp64" in res:
        log("device result before the host baseline: " + json.dumps(out))
        try:
            out["cpu_baseline"] = cpu_baseline(wl, res["fp64"]["stats"], args.cpu_threads, info["nnz_l"])
        except Exception as ex:  # noqa: BLE001 - the baseline is reported, never required
            log(f"cpu baseline failed: {ex}")
    if args.oracle_increments > 0:
        try:
            out["oracle_parity"] = oracle_parity(wl, res, args.oracle_increments, args.cpu_threads)
        except Exception as ex:  # noqa: BLE001 - reported, never required
            log(f"oracle parity failed: {ex}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
