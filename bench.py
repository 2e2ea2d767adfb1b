#!/usr/bin/env python3
"""Benchmark of the radiative solver on one MI355X.

A step is one solve of the narrow-band correlated-k OERM estimator over the
whole field: every emitting cell x photmax rays x nVar independent runs,
inputs (field, tables, grids) resident in HBM before the timed region.

Workload (default): 64x64x64 cells, x stretched (fact_mesh 1.5), synthetic
H2O-like tables of 360 bands x 16 g-points x 20 temperatures, photmax 200,
sorted bands (srt=1), nVar 8, FP32 march with FP64 tallies.

Legs: an untimed autotune of the batched schedule; the batched solver with
FP32 and with BF16 k-tables, each replayed as a hipGraph; the batched march
kernel alone (HIP events, no graph) for the roofline; the plain solver; one
solve at 4x photmax (checks eps ~ rays^-1/2, which the time-to-1%
projection assumes); the 256x128x256 field, batched against plain; the
correlated-k table generator (rc_ckdist) on 47 580 line-by-line spectra.

Printed (one JSON line on stdout): rays/s of the batched solver (FP32
k-tables unless --tables bf16), the plain solver's rays/s beside it, the
measured std-error eps of div(q_r) and the projected time to eps = 1% for
every leg, the roofline of the dominant kernel and a CPU baseline (the C++
oracle in FP32 on the host cores, a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
B_STEP = 32                     # 3 face coords + T + 2 k + 2 Ib texels (FP32)
B_RAY = 32                      # 2 x (2 k + 2 Ib texels) at emission and Tmax
B_STEP_BF16 = 28                # the same with the two k texels in BF16
B_RAY_BF16 = 28
B_CELL = 20                     # T in, FP64 tally pair out


RESTRICTION = 0                 # coarse-level T: 0 texture mean, 1 cubic splines


def new_session(p, tab, Ts):
    """A solver session with the fields Ts loaded (coarse levels by the
    chosen restriction)."""
    from radchan_amd import Solver
    s = Solver(p, tab)
    if RESTRICTION:
        s.set_restriction(RESTRICTION)
    s.set_fields(Ts)
    return s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def std_error(de: np.ndarray):
    """eps = sqrt(sum_c se_c^2) / sqrt(sum_c mean_c^2), se_c the standard
    error of the mean over the nVar runs; also the p95 of se_c/|mean_c| over
    cells with |mean_c| > 1% of the largest."""
    m = de.mean(axis=1)
    se = de.std(axis=1, ddof=1) / np.sqrt(de.shape[1])
    eps = float(np.sqrt((se ** 2).sum()) / np.sqrt((m ** 2).sum()))
    big = np.abs(m) > 0.01 * np.abs(m).max()
    p95 = float(np.percentile(se[big] / np.abs(m[big]), 95))
    return eps, p95


def cpu_baseline(p, tab, T, seconds: float):
    """The C++ oracle (FP32, reference semantics) on a sample of cells of the
    same workload, on the host cores."""
    from oracle import oracle as O
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(ncores, 16))
    o = O.Oracle(p, tab, T)
    cells = O.all_cells(p)
    rng = np.random.default_rng(123)
    order = rng.permutation(len(cells))
    n = 16 * ncores
    t0 = time.perf_counter()
    o.solve(cells[order[:n]], nthreads=ncores)
    dt = time.perf_counter() - t0
    n = int(min(len(cells), max(n, n * seconds / max(dt, 1e-3))))
    sample = cells[order[:n]]
    t0 = time.perf_counter()
    _, _, _, steps = o.solve(sample, nthreads=ncores)
    dt = time.perf_counter() - t0
    rays = n * p.photmax * p.nVar
    return {"value": rays / dt, "unit": "rays/s", "cores": ncores, "kind": "port",
            "sample": f"{n} of {len(cells)} cells (random), all {p.photmax} rays x {p.nVar} runs each: "
                      f"{rays} rays, {int(steps.sum())} steps in {dt:.1f} s",
            "steps_per_s": float(steps.sum() / dt)}


def timed_leg(s, lib, variant: int, steps: int, warmup: int):
    """warmup untimed solves, then `steps` timed ones (device-synchronised
    on both sides); returns (wall seconds per step, last stats, march ms
    of every step)."""
    for _ in range(warmup):
        s.solve(variant)
    lib.rc_device_synchronize()
    march = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = s.solve(variant)
        march.append(st["march_ms"])
    lib.rc_device_synchronize()
    return (time.perf_counter() - t0) / steps, st, march


def eps_scaling(p, tab, T, sched, eps, bf16: bool):
    """eps at 4x photmax on the same field (one solve): the projection to
    eps = 1% assumes eps ~ rays^(-1/2), so the ratio should be ~0.5.  (The
    time is not 4x: above 256 rays per cell the band keys are sorted in LDS
    rather than in registers.)"""
    from radchan_amd import VARIANT_BATCHED, Solver
    import copy
    q = copy.deepcopy(p)
    q.photmax = 4 * p.photmax
    s = new_session(q, tab, [T])
    s.set_table_precision(bf16)
    if sched:
        s.set_schedule(sched["wg_per_cu"], sched["threshold"], sched["steps_per_iter"])
    s.set_graph(False)
    s.solve(VARIANT_BATCHED)
    st = s.solve(VARIANT_BATCHED)
    e4, _ = std_error(s.tallies())
    s.close()
    return {"photmax": q.photmax, "eps": e4, "eps_ratio": e4 / eps, "expected_ratio": 0.5,
            "march_ms": st["march_ms"]}


def target_leg(lib, tab):
    """The 256 x 128 x 256 narrow-band field: batched (autotuned schedule,
    FP32 k tables, hipGraph) against the plain solver, rays/s and time to
    eps = 1%."""
    from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver, config
    from radchan_amd import fields as F
    p = config("big256")
    p.nVar = 8
    T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=0)
    s = new_session(p, tab, [T])
    sched = s.autotune()
    s.set_graph(True)
    wall, st, _ = timed_leg(s, lib, VARIANT_BATCHED, 2, 1)
    eps, _ = std_error(s.tallies())
    s.set_graph(False)
    pwall, sp, _ = timed_leg(s, lib, VARIANT_PLAIN, 1, 1)
    eps_p, _ = std_error(s.tallies())
    s.close()
    t1 = wall * (eps / 0.01) ** 2
    tp = pwall * (eps_p / 0.01) ** 2
    log(f"target 256x128x256: batched {1e3 * wall:.1f} ms, plain {1e3 * pwall:.1f} ms, "
        f"time-to-1% speed-up {tp / t1:.2f}x")
    return {"workload": f"narrow-band correlated-k OERM {p.imax}x{p.jmax}x{p.kmax}, photmax {p.photmax}, "
                        f"nVar {p.nVar}, FP32 k tables",
            "schedule": sched, "rays": int(st["rays"]), "steps": int(st["steps"]),
            "batched": {"ms_per_step": 1e3 * wall, "rays_per_s": st["rays"] / wall, "eps": eps,
                        "time_to_1pct_s": t1},
            "plain": {"ms_per_step": 1e3 * pwall, "rays_per_s": sp["rays"] / pwall, "eps": eps_p,
                      "time_to_1pct_s": tp},
            "time_to_1pct_speedup_vs_plain": tp / t1}


def ref_generator_time(d):
    """The reference generator's own correlatedKT + calcProbT (built from
    nbcorrk/ into oracle/_ref by make -C oracle ref; 2 temperatures at its
    compile-time sizes) timed on the first 2 temperatures of the workload,
    one core, or None if that build is absent."""
    import struct
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_ckdist")
    if not os.path.exists(exe):
        return None
    nB = d["nB"]
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "in.bin"), "wb") as f:
            f.write(struct.pack("<i", nB))
            f.write(np.asarray(d["temps"][:2], "<f8").tobytes())
            for s in range(2 * nB):
                b, e = d["off"][s], d["off"][s + 1]
                f.write(struct.pack("<idd", int(e - b), d["wvl"][s], d["wvr"][s]))
                f.write(np.asarray(d["wvn"][b:e], "<f8").tobytes())
                f.write(np.asarray(d["kvn"][b:e], "<f8").tobytes())
        os.mkdir(os.path.join(tmp, "errors"))
        t0 = time.perf_counter()
        subprocess.run([exe, "in.bin", "out.bin"], cwd=tmp, check=True, capture_output=True)
        dt = time.perf_counter() - t0
    return {"value": 2 * nB / dt, "unit": "spectra/s", "cores": 1, "kind": "reference",
            "sample": f"nbcorrk correlatedKT + calcProbT on the first 2 of the 52 temperatures ({2 * nB} spectra)"}


def ckdist_roofline(ck):
    """Roofline object of the generator kernel: algorithmic bytes per launch
    over its event-timed duration against 8 TB/s; HBM traffic from the
    committed FETCH_SIZE/WRITE_SIZE summary of the same workload."""
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_ckdist.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    ach = ck["algorithmic_bytes"] / (ck["kernel_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_ckdist<4>",
            "kernel_ms": ck["kernel_ms"], "algorithmic_bytes": ck["algorithmic_bytes"]}


def ckdist_leg(cpu: bool, reps: int = 5):
    """The correlated-k generator (rc_ckdist) on line-by-line spectra of the
    reference generator's layout: 10 cm^-1 bands over 150-9300 cm^-1 at
    dwv 0.01 (915 bands x 1000 points) at 52 temperatures, 20000 k^0.1 bins,
    16 g-points.  Kernel time from HIP events on the call's stream; the call
    time adds the host-to-device copy of the spectra.  CPU baseline: the
    reference generator's own correlatedKT + calcProbT built from nbcorrk/
    (one core), else the oracle's restatement on a 2000-spectrum sample."""
    from radchan_amd.spectra import ckdist, synthetic_lbl
    from radchan_amd.tables import GQ16_F64
    d = synthetic_lbl()
    ns, npts, nq = len(d["off"]) - 1, len(d["wvn"]), len(GQ16_F64)
    args = (d["off"], d["wvn"], d["kvn"], d["wvl"], d["wvr"], GQ16_F64)
    ckdist(*args)                                        # warm-up
    ks, calls = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        g = ckdist(*args)
        calls.append((time.perf_counter() - t0) * 1e3)
        ks.append(g["kernel_ms"])
    kms = float(np.mean(ks))
    # the table step (rc_ck_tables, calcProbT on the distributions; host)
    from radchan_amd.spectra import tables_from_distributions
    from radchan_amd.tables import WQ16_F64
    nB, nTt = d["nB"], len(d["temps"])
    t0 = time.perf_counter()
    tables_from_distributions(d["temps"], d["wvl"][:nB], d["wvr"][:nB], g["kq"].reshape(nTt, nB, nq),
                              g["kavg"].reshape(nTt, nB), WQ16_F64)
    table_ms = (time.perf_counter() - t0) * 1e3
    alg = npts * 16 + ns * (8 * (nq + 3) + 8 * 3)       # points in; kq, kavg, kmin, kmax out; edges, offsets
    out = {"spectra": ns, "points": npts, "n_pwrk": 20000, "nq": nq, "kernel_ms": round(kms, 4),
           "call_ms_with_h2d": round(float(np.mean(calls)), 2), "table_step_ms": round(table_ms, 2),
           "spectra_per_s": ns / (kms * 1e-3),
           "points_per_s": npts / (kms * 1e-3), "algorithmic_bytes": int(alg),
           "hbm_frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "cpu_baseline": None}
    if cpu:
        from oracle import oracle as O
        sub = 2000
        nth = min(os.cpu_count() or 1, 16)
        t0 = time.perf_counter()
        O.ckdist(d["off"][:sub + 1], d["wvn"], d["kvn"], d["wvl"][:sub], d["wvr"][:sub], GQ16_F64, nthreads=nth)
        dt = time.perf_counter() - t0
        port = {"value": sub / dt, "unit": "spectra/s", "cores": nth, "kind": "port",
                "sample": f"first {sub} of the {ns} spectra"}
        out["cpu_baseline"] = port
        ref = ref_generator_time(d)
        if ref is not None:
            out["cpu_baseline"] = ref
            out["cpu_port"] = port
        out["speedup_vs_cpu"] = out["spectra_per_s"] / out["cpu_baseline"]["value"]
    log(f"ckdist: {ns} spectra x {npts // ns} points in {kms:.3f} ms (kernel), "
        f"{out['call_ms_with_h2d']:.1f} ms with the copy in")
    return out


def bench_c4(args) -> int:
    """The snapshot ensemble: 8 snapshots of 64x96x64 and 4 of 256x128x256
    (laminar 300-1800 K profile plus 5% smoothed noise, its own noise seed
    per snapshot), photmax 100, nVar 8, FP32 k tables.  The batched solver
    runs all 12 snapshots, both grid sizes, in one launch (a group of the
    two sessions); the plain solver one launch per snapshot and run.  eps
    and the time to 1% are those of the worst large snapshot."""
    from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver, config
    from radchan_amd import fields as F
    from radchan_amd import lib as L
    from radchan_amd.tables import synthetic_nb

    tab = synthetic_nb()
    lib = L.load()
    shapes = []
    for name, n, seed0 in (("big256", 4, 100), ("ens64x96", 8, 200)):
        p = config(name)
        p.nVar = args.nvar
        Ts = [F.channel_field(p, 300.0, 1800.0, noise=0.05, seed=seed0 + i) for i in range(n)]
        s = new_session(p, tab, Ts)
        shapes.append((name, p, n, s))
        log(f"c4: {n} x {p.imax}x{p.jmax}x{p.kmax}")
    from radchan_amd import Group
    group = Group([s for _, _, _, s in shapes])
    # the schedule the autotuner picks for a single 256x128x256 field
    group.set_schedule(32, 8, 4)

    def step(variant):
        if variant == VARIANT_BATCHED:
            return [group.solve()]
        for _, _, _, s in shapes:
            s.run(variant)
        return [s.sync() for _, _, _, s in shapes]

    def timed(variant, steps, warmup):
        for _ in range(warmup):
            step(variant)
        lib.rc_device_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sts = step(variant)
        lib.rc_device_synchronize()
        return (time.perf_counter() - t0) / steps, sts

    wall, sts = timed(VARIANT_BATCHED, args.steps, args.warmup)
    rays = sum(st["rays"] for st in sts)
    eps_l = [std_error(shapes[0][3].tallies(i))[0] for i in range(shapes[0][2])]
    log(f"c4 batched: {1e3 * wall:.1f} ms/step, {rays / wall / 1e9:.3f} Grays/s, eps (large) {max(eps_l):.4g}")
    plain = None
    if not args.no_plain:
        pwall, psts = timed(VARIANT_PLAIN, 1, 0)
        peps = [std_error(shapes[0][3].tallies(i))[0] for i in range(shapes[0][2])]
        plain = {"ms_per_step": 1e3 * pwall, "rays_per_s": sum(st["rays"] for st in psts) / pwall,
                 "eps_large_max": max(peps), "time_to_1pct_s": pwall * (max(peps) / 0.01) ** 2}
        log(f"c4 plain: {1e3 * pwall:.1f} ms/step")
    t1 = wall * (max(eps_l) / 0.01) ** 2
    out = {
        "metric": "rays/s (batched solver) and time to 1% std-error of div q_r; batched vs plain, 1 MI355X",
        "value": rays / wall, "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 march, fp32 k-tables, fp64 tallies",
        "data": "synthetic (H2O-like tables, laminar 300-1800 K channel T fields with 5% smoothed noise)",
        "config": {"workload": "snapshot ensemble: 8 x 64x96x64 + 4 x 256x128x256, photmax 100, nVar %d" % args.nvar,
                   "rays_per_step": int(rays), "batched_launches_per_step": 1,
                   "schedule": {"wg_per_cu": 32, "threshold": 8, "steps_per_iter": 4}},
        "eps_large": eps_l, "time_to_1pct_s": t1,
        "plain": plain,
        "time_to_1pct_speedup_vs_plain": (plain["time_to_1pct_s"] / t1) if plain else None,
    }
    print(json.dumps(out), flush=True)
    group.close()
    for _, _, _, s in shapes:
        s.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="nb64", choices=["nb64", "nb128", "big256", "grey32", "mg192", "c4",
                                                               "ckdist"])
    ap.add_argument("--nvar", type=int, default=8)
    ap.add_argument("--tables", default="fp32", choices=["bf16", "fp32"],
                    help="k-table storage of the headline (batched) leg; both are measured")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-ckdist", action="store_true", help="skip the correlated-k generator leg")
    ap.add_argument("--restriction", default="texture", choices=["texture", "spline"],
                    help="coarse-level temperatures of multigrid workloads: texture mean (as the reference) or "
                         "cubic splines (interp3D)")
    ap.add_argument("--no-target", action="store_true",
                    help="skip the 256x128x256 batched-vs-plain leg")
    ap.add_argument("--probe", action="store_true",
                    help="no timing: one elementwise copy of 512 MiB (counter calibration) and one batched "
                         "solve, for hardware-counter runs")
    args = ap.parse_args()
    global RESTRICTION
    RESTRICTION = 1 if args.restriction == "spline" else 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU path (cell slabs per GPU) is not "
              "part of it, so --gpus must be 1", file=sys.stderr)
        return 2

    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "ckdist":
        ck = ckdist_leg(not args.no_cpu, reps=args.steps)
        print(json.dumps({"metric": "correlated-k distributions per second (rc_ckdist), 1 MI355X",
                          "value": ck["spectra_per_s"], "unit": "spectra/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": 1, "ms_per_step": ck["kernel_ms"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic line spectra",
                          "config": {"workload": f"{ck['spectra']} spectra x {ck['points'] // ck['spectra']} points, "
                                                 f"n_pwrk {ck['n_pwrk']}, nq {ck['nq']}"},
                          "roofline": ckdist_roofline(ck),
                          "ckdist": ck, "cpu_baseline": ck["cpu_baseline"]}), flush=True)
        return 0

    from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver, config
    from radchan_amd import fields as F
    from radchan_amd import lib as L
    from radchan_amd.tables import synthetic_nb

    p = config(args.workload)
    p.nVar = args.nvar
    tab = None
    if p.grey:
        T = F.channel_field(p, 500.0, 1000.0, noise=0.02, seed=0)
    else:
        tab = synthetic_nb()
        T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=0)
    log(f"bench: workload {args.workload} {p.imax}x{p.jmax}x{p.kmax} photmax {p.photmax} nVar {p.nVar}")
    s = new_session(p, tab, [T])
    lib = L.load()
    bf16 = args.tables == "bf16" and not p.grey

    if args.probe:
        # a device-to-device copy of 512 MiB through the HIP runtime (a blit
        # kernel: reads and writes 512 MiB), then one batched solve
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = 512 << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
            raise RuntimeError("hipMalloc failed")
        hip.hipMemset(a, 1, ctypes.c_size_t(n))
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)           # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        s.set_table_precision(bf16)
        st = s.solve(VARIANT_BATCHED)
        hip.hipFree(a)
        hip.hipFree(b)
        print(json.dumps({"probe": args.workload, "tables": args.tables, "copy_bytes": n,
                          "rays": st["rays"], "steps": st["steps"], "march_ms": st["march_ms"]}), flush=True)
        s.close()
        return 0

    sched = None
    if not args.no_autotune:
        sched = s.autotune()                       # untimed: picks wavefronts per CU and refill threshold
        log(f"autotune: {sched}")

    legs = {}
    for name, use_bf16 in (("fp32_tables", False), ("bf16_tables", True)):
        if use_bf16 and p.grey:
            continue
        s.set_table_precision(use_bf16)
        s.set_graph(not args.no_graph)
        wall, st, _ = timed_leg(s, lib, VARIANT_BATCHED, args.steps, args.warmup)
        eps, p95 = std_error(s.tallies())
        legs[name] = {"rays_per_s": st["rays"] / wall, "ms_per_step": 1e3 * wall, "eps": eps, "eps_p95_cell": p95,
                      "time_to_1pct_s": wall * (eps / 0.01) ** 2, "rays": st["rays"], "steps": st["steps"]}
        log(f"batched {name}: {1e3 * wall:.2f} ms/step, {st['rays'] / wall / 1e9:.3f} Grays/s, "
            f"{st['steps'] / st['rays']:.2f} steps/ray, eps {eps:.4g}")
    head = legs["bf16_tables" if bf16 else "fp32_tables"]

    # the march kernel alone, outside a graph: HIP events on the solver's stream
    s.set_table_precision(bf16)
    s.set_graph(False)
    _, st, march = timed_leg(s, lib, VARIANT_BATCHED, max(3, args.steps // 2), 1)
    rays, steps = st["rays"], st["steps"]

    plain = None
    if not args.no_plain and args.plain_steps > 0:
        s.set_table_precision(False)
        pwall, sp, _ = timed_leg(s, lib, VARIANT_PLAIN, args.plain_steps, 1)
        eps_p, _ = std_error(s.tallies())
        plain = {"rays_per_s": sp["rays"] / pwall, "ms_per_step": 1e3 * pwall, "eps": eps_p,
                 "time_to_1pct_s": pwall * (eps_p / 0.01) ** 2}
        log(f"plain:   {1e3 * pwall:.2f} ms/step, {sp['rays'] / pwall / 1e9:.3f} Grays/s, eps {eps_p:.4g}")

    ncell = s.ncell
    b_step, b_ray = (B_STEP_BF16, B_RAY_BF16) if bf16 else (B_STEP, B_RAY)
    alg_bytes = steps * b_step + rays * b_ray + ncell * B_CELL
    march_ms = float(np.mean(march))
    achieved = alg_bytes / (march_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.tables}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "%s<false,%d> (persistent march, %s k-tables)" % (
                "k_batched_mg" if p.grid_num > 1 else "k_batched", 1 if bf16 else 0, "BF16" if bf16 else "FP32"),
            "kernel_ms": round(march_ms, 4),
            "algorithmic_bytes": int(alg_bytes),
            "bytes_per_step": b_step, "bytes_per_ray": b_ray, "bytes_per_cell": B_CELL}

    scal = eps_scaling(p, tab, T, sched, head["eps"], bf16)
    log(f"eps at photmax {scal['photmax']}: {scal['eps']:.4g} (ratio {scal['eps_ratio']:.3f}, expected 0.5), "
        f"{scal['march_ms']:.1f} ms")
    target = None
    if not args.no_target and args.workload != "big256" and not p.grey:
        target = target_leg(lib, tab)

    ck = None if args.no_ckdist else ckdist_leg(not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(p, tab, T, args.cpu_seconds)
            log(f"cpu:     {cpu['value'] / 1e6:.2f} Mrays/s on {cpu['cores']} cores")
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {e}")

    value = head["rays_per_s"]
    out = {
        "metric": "rays/s (batched solver) and time to 1% std-error of div q_r; batched vs plain, 1 MI355X",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 march, bf16 k-tables, fp64 tallies" if bf16 else "fp32 march, fp32 k-tables, fp64 tallies",
        "data": "synthetic (H2O-like tables, parabolic channel T field with 2% smoothed noise)",
        "config": {"workload": f"narrow-band correlated-k OERM {p.imax}x{p.jmax}x{p.kmax}, {p.nB} bands x "
                               f"{p.nQ} g x {p.nT} T, photmax {p.photmax}, nVar {p.nVar}, srt {p.srt}, "
                               f"grid_num {p.grid_num}" if not p.grey else
                               f"grey OERM {p.imax}x{p.jmax}x{p.kmax}, photmax {p.photmax}, nVar {p.nVar}",
                   "cells": ncell, "rays_per_step": int(rays), "steps_per_ray": steps / rays,
                   "schedule": sched, "graph": not args.no_graph, "restriction": args.restriction},
        "eps": head["eps"],
        "eps_p95_cell": head["eps_p95_cell"],
        "time_to_1pct_s": head["time_to_1pct_s"],
        "batched": legs,
        "plain": plain,
        "speedup_vs_plain": (value / plain["rays_per_s"]) if plain else None,
        "time_to_1pct_speedup_vs_plain": (plain["time_to_1pct_s"] / head["time_to_1pct_s"]) if plain else None,
        "eps_scaling": scal,
        "target_256x128x256": target,
        "ckdist": ck,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
