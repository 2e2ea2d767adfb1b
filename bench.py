#!/usr/bin/env python3
"""Benchmark of the radiative solver on one MI355X.

A step is one solve of the narrow-band correlated-k OERM estimator over the
whole field: every emitting cell x photmax rays x nVar independent runs,
inputs (field, tables, grids) resident in HBM before the timed region.

Workload (default, the headline): the 256x128x256 narrow-band field, x
stretched (fact_mesh 1.5), synthetic H2O-like tables of 360 bands x 16
g-points x 20 temperatures, photmax 100, sorted bands (srt=1), nVar 8, FP32
march with FP64 tallies.

Legs on the headline field: an untimed autotune of the batched schedule;
the batched solver with FP32 and with BF16 k-tables, each replayed as a
hipGraph; the batched march kernel alone (HIP events, no graph) for the
roofline; the plain solver (the reference's launch shape, one launch per
run); one solve at 4x photmax (checks eps ~ rays^-1/2, which the time-to-1%
projection assumes).  Beside it: the 64^3 field (photmax 200) with the same
legs and a measured run at the projected budget for eps = 1% (one launch
of as many runs as the projection asks); the correlated-k table generator
(rc_ckdist) on 47 580 line-by-line spectra; the CPU baseline (the C++
oracle in FP32 on the host's CPU share and on one core, bounded samples of
the grey 32^3, the 64^3 and the 256x128x256 fields).

Printed (one JSON line on stdout): rays/s of the batched solver on the
headline field (FP32 k-tables unless --tables bf16), the plain solver's
rays/s beside it, the measured std-error eps of div(q_r) and the projected
time to eps = 1% for every leg, the roofline of the dominant kernel and the
CPU baseline.
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


def host_cpu() -> dict:
    """The host's CPU: model name (lscpu, else /proc/cpuinfo), nproc, the
    CPUs this process may run on, and the share the baseline uses: those
    CPUs, capped by OMP_NUM_THREADS when set (the GPU box gives one GPU's
    job a share of its cores that way)."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    share = max(1, min(aff, omp) if omp > 0 else aff)
    return {"model": model, "nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": omp or None,
            "threads_used": share}


def cpu_sample(o, p, cells, seconds: float, threads: int, seed: int = 123) -> dict:
    """The oracle (FP32, reference semantics) on a random sample of the
    listed cells, sized to about `seconds` of work at `threads` threads:
    rays/s, ray-steps/s and the wall time of the sample."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(cells))
    n = min(len(cells), max(4, 4 * threads))
    t0 = time.perf_counter()
    o.solve(cells[order[:n]], nthreads=threads)
    dt = time.perf_counter() - t0
    n = int(min(len(cells), max(n, n * seconds / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, _, _, steps = o.solve(cells[order[:n]], nthreads=threads)
    dt = time.perf_counter() - t0
    rays = n * p.photmax * p.nVar
    return {"value": rays / dt, "unit": "rays/s", "steps_per_s": float(steps.sum() / dt), "cores": threads,
            "wall_s": dt, "sample": f"{n} of {len(cells)} cells (random), all {p.photmax} rays x {p.nVar} runs "
                                    f"each: {rays} rays, {int(steps.sum())} steps"}


def cpu_baseline(p, tab, T, seconds: float, host: dict) -> dict:
    """The C++ oracle on the host: the CPU share and one core, each on a
    bounded sample of the same workload."""
    from oracle import oracle as O
    o = O.Oracle(p, tab, T)
    cells = O.all_cells(p)
    share = cpu_sample(o, p, cells, seconds, host["threads_used"])
    one = cpu_sample(o, p, cells, seconds, 1)
    o.close()
    out = dict(share)
    out.update({"kind": "port", "host": host, "one_core": one,
                "scaling_vs_one_core": share["value"] / one["value"]})
    return out


def cpu_configs(tab, seconds: float, host: dict) -> dict:
    """The oracle on the grey 32^3 and narrow-band 64^3 fields (the CPU
    share, bounded samples), beside the headline field's baseline."""
    from oracle import oracle as O
    from radchan_amd import config
    from radchan_amd import fields as F
    out = {}
    for name in ("grey32", "nb64"):
        p = config(name)
        p.nVar = 8
        if p.grey:
            T = F.channel_field(p, 500.0, 1000.0, noise=0.02, seed=0)
            o = O.Oracle(p, None, T)
        else:
            T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=0)
            o = O.Oracle(p, tab, T)
        out[name] = cpu_sample(o, p, O.all_cells(p), seconds, host["threads_used"])
        o.close()
    return out


def lane_use(st) -> dict:
    """Lane use of the batched march: the fraction of the bulk march's lane
    slots (64 per wavefront sub-step) that stepped a live ray, and the
    fraction of all steps taken in the drain (after a wavefront has no unit
    left to start)."""
    slots, bulk = st.get("bulk_slots", 0), st.get("bulk_steps", 0)
    return {"bulk_active_lane_frac": bulk / slots if slots else None,
            "drain_step_frac": 1.0 - bulk / st["steps"] if st["steps"] else None}


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
    """eps at 4x photmax on the same field (one solve; 2x where 4x would pass
    the 16384 rays per cell that the band sort holds): the projection to
    eps = 1% assumes eps ~ rays^(-1/2), so the ratio should be ~mult^(-1/2).
    (The time is not exactly 4x: above 256 rays per cell the band keys are
    sorted by radix passes through HBM rather than in registers.)"""
    from radchan_amd import VARIANT_BATCHED
    import copy
    mult = next((m for m in (4, 2) if p.srt != 1 or m * p.photmax <= 16384), None)
    if mult is None:
        return None
    q = copy.deepcopy(p)
    q.photmax = mult * p.photmax
    s = new_session(q, tab, [T])
    s.set_table_precision(bf16)
    if sched:
        s.set_schedule(sched["wg_per_cu"], sched["threshold"], sched["steps_per_iter"], sched.get("refill", -1))
    s.set_graph(False)
    s.solve(VARIANT_BATCHED)
    st = s.solve(VARIANT_BATCHED)
    e4, _ = std_error(s.tallies())
    s.close()
    return {"photmax": q.photmax, "eps": e4, "eps_ratio": e4 / eps, "expected_ratio": round(mult ** -0.5, 4),
            "march_ms": st["march_ms"]}


def field_legs(args, lib, p, tab, T, bf16: bool, steps: int, warmup: int, plain_steps: int, scaling: bool = True,
               fp64_steps: int = 0):
    """Every leg of one field: autotune, the batched solver with FP32 and
    BF16 k-tables (hipGraph replays, `steps` timed), the march kernel alone
    (HIP events on the solver's stream, no graph) for the roofline, the
    plain solver, and (scaling) one solve at 4x photmax."""
    from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN
    s = new_session(p, tab, [T])
    sched = None
    if not args.no_autotune:
        sched = s.autotune()                       # untimed: picks the schedule and refill mode
        log(f"autotune {p.imax}x{p.jmax}x{p.kmax}: {sched}")
    legs = {}
    for name, use_bf16 in (("fp32_tables", False), ("bf16_tables", True)):
        if use_bf16 and p.grey:
            continue
        s.set_table_precision(use_bf16)
        s.set_graph(not args.no_graph)
        wall, st, _ = timed_leg(s, lib, VARIANT_BATCHED, steps, warmup)
        eps, p95 = std_error(s.tallies())
        legs[name] = {"rays_per_s": st["rays"] / wall, "ms_per_step": 1e3 * wall, "eps": eps, "eps_p95_cell": p95,
                      "time_to_1pct_s": wall * (eps / 0.01) ** 2, "rays": st["rays"], "steps": st["steps"],
                      "timed_solves": steps, **lane_use(st)}
        log(f"batched {name}: {1e3 * wall:.2f} ms/step, {st['rays'] / wall / 1e9:.3f} Grays/s, "
            f"{st['steps'] / st['rays']:.2f} steps/ray, eps {eps:.4g}, bulk active lanes "
            f"{legs[name]['bulk_active_lane_frac']:.3f}, drain steps {legs[name]['drain_step_frac']:.4f}")
    head = legs["bf16_tables" if bf16 else "fp32_tables"]

    # the march kernel alone, outside a graph: HIP events on the solver's stream
    s.set_table_precision(bf16)
    s.set_graph(False)
    _, st, march = timed_leg(s, lib, VARIANT_BATCHED, max(3, steps), 1)
    rays, nsteps = st["rays"], st["steps"]

    det = None
    if not args.no_deterministic:
        # the deterministic reduction: per-ray tallies in HBM, summed in a fixed order
        s.set_reduction(1)
        dwall, sd, _ = timed_leg(s, lib, VARIANT_BATCHED, 2, 1)
        s.set_reduction(0)
        det = {"ms_per_step": 1e3 * dwall, "rays_per_s": sd["rays"] / dwall,
               "overhead_vs_atomic": dwall * 1e3 / head["ms_per_step"] - 1.0, "timed_solves": 2,
               "ray_tally_bytes": int(sd["rays"]) * 4}
        log(f"batched, deterministic reduction: {1e3 * dwall:.2f} ms/step")

    sobol = None
    if p.sampler == 0 and not p.grey and not args.no_secondary:
        # param.h's default direction sampler (random 1): Sobol points shifted per (cell, run)
        q = p.copy()
        q.sampler = 1
        s1 = new_session(q, tab, [T])
        s1.set_table_precision(bf16)
        if sched:
            s1.set_schedule(sched["wg_per_cu"], sched["threshold"], sched["steps_per_iter"], sched.get("refill", -1))
        s1.set_graph(not args.no_graph)
        swall, ss, _ = timed_leg(s1, lib, VARIANT_BATCHED, 2, 1)
        eps_s, _ = std_error(s1.tallies())
        s1.close()
        sobol = {"ms_per_step": 1e3 * swall, "rays_per_s": ss["rays"] / swall, "eps": eps_s,
                 "time_to_1pct_s": swall * (eps_s / 0.01) ** 2, "eps_vs_philox": eps_s / head["eps"], "timed_solves": 2}
        log(f"batched, Sobol directions: {1e3 * swall:.2f} ms/step, eps {eps_s:.4g} ({eps_s / head['eps']:.3f} of Philox)")

    plain = None
    if not args.no_plain and plain_steps > 0:
        s.set_table_precision(False)
        pwall, sp, _ = timed_leg(s, lib, VARIANT_PLAIN, plain_steps, 1)
        eps_p, _ = std_error(s.tallies())
        plain = {"rays_per_s": sp["rays"] / pwall, "ms_per_step": 1e3 * pwall, "eps": eps_p,
                 "time_to_1pct_s": pwall * (eps_p / 0.01) ** 2, "timed_solves": plain_steps}
        log(f"plain:   {1e3 * pwall:.2f} ms/step, {sp['rays'] / pwall / 1e9:.3f} Grays/s, eps {eps_p:.4g}")
    bands = None
    if not args.no_plain and args.plain_bands_steps > 0 and p.srt == 1 and not p.grey:
        from radchan_amd import VARIANT_PLAIN_BANDS
        s.set_table_precision(False)
        bwall, sb, _ = timed_leg(s, lib, VARIANT_PLAIN_BANDS, args.plain_bands_steps, 0)
        eps_b, _ = std_error(s.tallies())
        bands = {"rays_per_s": sb["rays"] / bwall, "ms_per_step": 1e3 * bwall, "eps": eps_b,
                 "time_to_1pct_s": bwall * (eps_b / 0.01) ** 2, "timed_solves": args.plain_bands_steps,
                 "launches_per_step": sb["launches"]}
        log(f"plain, one launch per band: {1e3 * bwall:.2f} ms/step ({sb['launches']} launches), eps {eps_b:.4g}")
    ncell = s.ncell
    s.close()
    fp64 = None
    if not args.no_plain and fp64_steps > 0 and plain:
        # floatpoint 1: the full-FP64 kernels in the plain launch shape
        q = p.copy()
        q.fp64 = 1
        s64 = new_session(q, tab, [T])
        fwall, sf, _ = timed_leg(s64, lib, VARIANT_PLAIN, fp64_steps, 1)
        eps_f, _ = std_error(s64.tallies())
        s64.close()
        fp64 = {"rays_per_s": sf["rays"] / fwall, "ms_per_step": 1e3 * fwall, "eps": eps_f, "timed_solves": fp64_steps,
                "time_vs_fp32_plain": fwall * 1e3 / plain["ms_per_step"]}
        log(f"plain, FP64 (floatpoint 1): {1e3 * fwall:.2f} ms/step, {fp64['time_vs_fp32_plain']:.2f}x the FP32 plain")

    b_step, b_ray = (B_STEP_BF16, B_RAY_BF16) if bf16 else (B_STEP, B_RAY)
    alg_bytes = nsteps * b_step + rays * b_ray + ncell * B_CELL
    march_ms = float(np.mean(march))
    achieved = alg_bytes / (march_ms * 1e-3) / 1e9
    wl = {(64, 64, 64): "nb64", (256, 128, 256): "big256", (192, 128, 192): "mg192"}.get((p.imax, p.jmax, p.kmax))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{wl}_{'bf16' if bf16 else 'fp32'}.json")
    if wl and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "%s<false, %d, %s> (persistent march, %s k-tables)" % (
                "k_batched_mg" if p.grid_num > 1 else "k_batched", 1 if bf16 else 0,
                "true" if (p.srt == 1 and not p.grey and p.photmax > 256) else "false", "BF16" if bf16 else "FP32"),
            "kernel_ms": round(march_ms, 4), "kernel_ms_each": [round(float(m), 4) for m in march],
            "algorithmic_bytes": int(alg_bytes), "steps_per_launch": int(nsteps), "rays_per_launch": int(rays),
            "bytes_per_step": b_step, "bytes_per_ray": b_ray, "bytes_per_cell": B_CELL}
    out = {"workload": f"narrow-band correlated-k OERM {p.imax}x{p.jmax}x{p.kmax}, {p.nB} bands x {p.nQ} g x "
                       f"{p.nT} T, photmax {p.photmax}, nVar {p.nVar}, srt {p.srt}, grid_num {p.grid_num}"
                       if not p.grey else f"grey OERM {p.imax}x{p.jmax}x{p.kmax}, photmax {p.photmax}, nVar {p.nVar}",
           "cells": ncell, "schedule": sched, "head": head, "batched": legs, "plain": plain,
           "plain_one_launch_per_band": bands, "plain_fp64": fp64, "deterministic_reduction": det,
           "sobol_directions": sobol, "roofline": roof,
           "speedup_vs_plain": (head["rays_per_s"] / plain["rays_per_s"]) if plain else None,
           "time_to_1pct_speedup_vs_plain": (plain["time_to_1pct_s"] / head["time_to_1pct_s"]) if plain else None,
           "time_to_1pct_speedup_vs_plain_per_band": (bands["time_to_1pct_s"] / head["time_to_1pct_s"])
           if bands else None}
    if scaling:
        sc = eps_scaling(p, tab, T, sched, head["eps"], bf16)
        if sc:
            log(f"eps at photmax {sc['photmax']}: {sc['eps']:.4g} (ratio {sc['eps_ratio']:.3f}, expected "
                f"{sc['expected_ratio']}), {sc['march_ms']:.1f} ms")
        out["eps_scaling"] = sc
    return out


def confirm_leg(lib, p, tab, T, sched, eps: float, ms: float):
    """The projected time to eps = 1% measured: one batched launch of as
    many independent runs per cell as the projection from `eps` (nVar
    runs per solve) asks, plus 10% (the eps of one solve is itself an
    estimate); eps of the result from the per-cell means and the
    reference's sqrt-variance (the standard error of the mean is
    sdev sqrt((n - 1/2) / n))."""
    import copy
    import math
    from radchan_amd import VARIANT_BATCHED
    q = copy.deepcopy(p)
    q.nVar = int(math.ceil(p.nVar * (eps / 0.01) ** 2 * 1.1))
    s = new_session(q, tab, [T])
    if sched:
        s.set_schedule(sched["wg_per_cu"], sched["threshold"], sched["steps_per_iter"], sched.get("refill", -1))
    s.set_graph(False)
    lib.rc_device_synchronize()
    t0 = time.perf_counter()
    st = s.solve(VARIANT_BATCHED)
    lib.rc_device_synchronize()
    wall = time.perf_counter() - t0
    mean, sd = s.results()
    s.close()
    n = q.nVar
    m = mean.astype(np.float64).reshape(-1)
    se = sd.astype(np.float64).reshape(-1) * math.sqrt((n - 0.5) / n)
    eps_m = float(np.sqrt((se ** 2).sum()) / np.sqrt((m ** 2).sum()))
    proj = ms * 1e-3 * (eps / 0.01) ** 2
    log(f"time to 1%: {q.nVar} runs per cell in one launch, {wall:.2f} s, eps {eps_m:.5f} (projected {proj:.2f} s)")
    return {"runs_per_cell": n, "rays": int(st["rays"]), "steps": int(st["steps"]), "wall_s": wall,
            "kernel_s": st["march_ms"] * 1e-3, "eps": eps_m, "projected_s": proj,
            "measured_over_projected": wall / proj}


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
    group.set_schedule(32, 4, 8, 1)

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
    # the mixed-precision form: BF16 k tables in both sessions, one group launch
    group.close()
    for _, _, _, s in shapes:
        s.set_table_precision(True)
    group = Group([s for _, _, _, s in shapes])
    group.set_schedule(32, 4, 8, 1)
    bwall, _ = timed(VARIANT_BATCHED, args.steps, args.warmup)
    beps = [std_error(shapes[0][3].tallies(i))[0] for i in range(shapes[0][2])]
    bf16 = {"ms_per_step": 1e3 * bwall, "rays_per_s": rays / bwall, "eps_large_max": max(beps),
            "time_to_1pct_s": bwall * (max(beps) / 0.01) ** 2}
    log(f"c4 batched, BF16 tables: {1e3 * bwall:.1f} ms/step")
    group.close()
    for _, _, _, s in shapes:
        s.set_table_precision(False)
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
                   "schedule": {"wg_per_cu": 32, "threshold": 4, "steps_per_iter": 8, "refill": 1}},
        "eps_large": eps_l, "time_to_1pct_s": t1,
        "batched_bf16_tables": bf16,
        "plain": plain,
        "time_to_1pct_speedup_vs_plain": (plain["time_to_1pct_s"] / t1) if plain else None,
    }
    print(json.dumps(out), flush=True)
    group.close()
    for _, _, _, s in shapes:
        s.close()
    return 0


def bench_slab(args) -> int:
    """One GPU's share of param.h's production run as it stands: the whole
    406x576x2304 field resident (a laminar 300-1500 K profile with a 3%
    sin/cos modulation), the 48-plane z slab of rank 24 of p_col = 48 solved
    (406x576x48 cells), photmax 6000, nine levels with maxs 5, bd 4 z walls,
    Sobol directions, NarrowBand.h's table sizes, nVar 2 (param.h has 1; two
    runs give the std-error).  Legs: the batched solver at its default
    schedule (no autotune: a solve takes seconds), BF16 tables, and the
    plain solver, each timed once after (batched) one warm-up solve."""
    from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver, config
    from radchan_amd import fields as F
    from radchan_amd import lib as L
    from radchan_amd.tables import synthetic_nb

    p = config("slab406")
    p.nVar = args.nvar if args.nvar_given else 2
    rank = p.p_col // 2
    kloc = p.kmax // p.p_col
    z0 = rank * kloc + 1
    lib = L.load()
    log(f"bench: workload slab406 {p.imax}x{p.jmax}x{p.kmax} field, z slab {z0}..{z0 + kloc - 1} of rank {rank} "
        f"(p_col {p.p_col}), photmax {p.photmax}, nVar {p.nVar}")
    t0 = time.perf_counter()
    T = F.wavy_channel_field(p, 300.0, 1500.0, amp=0.03)
    tab = synthetic_nb(p.nB, p.nT, p.nQ)
    log(f"field and tables built in {time.perf_counter() - t0:.1f} s")
    s = Solver(p, tab)
    s.set_fields([T], zstart=z0)
    # the library's defaults on several levels (refills once per iteration,
    # steps per iteration = the finest level's cap, maxs 5: the fastest of
    # the schedules timed on this share), set here so that the line names
    # what ran
    sched = {"wg_per_cu": 20, "threshold": 4, "steps_per_iter": p.maxs[0], "refill": 0}
    s.set_schedule(sched["wg_per_cu"], sched["threshold"], sched["steps_per_iter"], sched["refill"])
    if args.probe:
        # no timing: one batched solve, for hardware-counter runs
        st = s.solve(VARIANT_BATCHED)
        print(json.dumps({"probe": "slab406", "nVar": p.nVar, "rays": st["rays"], "steps": st["steps"],
                          "march_ms": st["march_ms"], "schedule": sched}), flush=True)
        s.close()
        return 0
    steps = max(1, min(args.steps, 2))
    legs = {}
    for name, use_bf16 in (("fp32_tables", False), ("bf16_tables", True)):
        s.set_table_precision(use_bf16)
        wall, st, march = timed_leg(s, lib, VARIANT_BATCHED, steps, 1)
        eps, p95 = std_error(s.tallies())
        legs[name] = {"rays_per_s": st["rays"] / wall, "ms_per_step": 1e3 * wall, "march_ms": march, "eps": eps,
                      "eps_p95_cell": p95, "time_to_1pct_s": wall * (eps / 0.01) ** 2, "rays": st["rays"],
                      "steps": st["steps"], "timed_solves": steps, **lane_use(st)}
        log(f"batched {name}: {1e3 * wall:.1f} ms/step, {st['rays'] / wall / 1e9:.3f} Grays/s, "
            f"{st['steps'] / st['rays']:.2f} steps/ray, eps {eps:.4g}, bulk active lanes "
            f"{legs[name]['bulk_active_lane_frac']:.3f}")
    head = legs["fp32_tables"]
    s.set_table_precision(False)
    plain = None
    if not args.no_plain:
        pwall, sp, _ = timed_leg(s, lib, VARIANT_PLAIN, 1, 0)
        eps_p, _ = std_error(s.tallies())
        plain = {"rays_per_s": sp["rays"] / pwall, "ms_per_step": 1e3 * pwall, "eps": eps_p,
                 "time_to_1pct_s": pwall * (eps_p / 0.01) ** 2, "timed_solves": 1}
        log(f"plain:   {1e3 * pwall:.1f} ms/step, {sp['rays'] / pwall / 1e9:.3f} Grays/s, eps {eps_p:.4g}")
    ncell = s.ncell
    s.close()
    march_ms = float(np.mean(head["march_ms"]))
    alg_bytes = head["steps"] * B_STEP + head["rays"] * B_RAY + ncell * B_CELL
    achieved = alg_bytes / (march_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_batched_mg<false, 0, true> (persistent march, FP32 k-tables)",
            "kernel_ms": round(march_ms, 4), "algorithmic_bytes": int(alg_bytes),
            "steps_per_launch": int(head["steps"]), "rays_per_launch": int(head["rays"]),
            "bytes_per_step": B_STEP, "bytes_per_ray": B_RAY, "bytes_per_cell": B_CELL}
    out = {
        "metric": "rays/s (batched solver) and time to 1% std-error of div q_r; batched vs plain, 1 MI355X",
        "value": head["rays_per_s"], "unit": "rays/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 march, fp32 k-tables, fp64 tallies",
        "data": "synthetic (H2O-like tables of NarrowBand.h's sizes, laminar 300-1500 K channel T field, 3% "
                "sin/cos modulation)",
        "config": {"workload": f"param.h production share: {p.imax}x{p.jmax}x{p.kmax} field, z slab of rank {rank} "
                               f"of p_col {p.p_col} ({p.imax}x{p.jmax}x{kloc} cells), photmax {p.photmax}, nVar "
                               f"{p.nVar}, grid_num {p.grid_num}, Sobol directions, {p.nB} bands x {p.nQ} g x "
                               f"{p.nT} T", "zstart": z0, "cells": ncell, "rays_per_step": int(head["rays"]),
                   "steps_per_ray": head["steps"] / head["rays"], "schedule": sched},
        "eps": head["eps"], "time_to_1pct_s": head["time_to_1pct_s"],
        "batched": legs, "plain": plain,
        "speedup_vs_plain": (head["rays_per_s"] / plain["rays_per_s"]) if plain else None,
        "roofline": roof,
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="big256", choices=["nb64", "nb128", "big256", "grey32", "mg192", "prod28",
                                                                 "c4", "ckdist", "slab406"])
    ap.add_argument("--nvar", type=int, default=None, help="runs per cell (default 8; slab406: 2)")
    ap.add_argument("--tables", default="fp32", choices=["bf16", "fp32"],
                    help="k-table storage of the headline (batched) leg; both are measured")
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="timed solves of the plain solver (default 2 on fields of 4 M cells or more, else 3)")
    ap.add_argument("--plain-bands-steps", type=int, default=1,
                    help="timed solves of the plain solver launched once per sorted band (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-deterministic", action="store_true",
                    help="skip the batched solve with the deterministic reduction")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-ckdist", action="store_true", help="skip the correlated-k generator leg")
    ap.add_argument("--restriction", default="texture", choices=["texture", "spline"],
                    help="coarse-level temperatures of multigrid workloads: texture mean (as the reference) or "
                         "cubic splines (interp3D)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the 64^3 legs and the measured time to 1%% beside the headline")
    ap.add_argument("--no-confirm", action="store_true",
                    help="--workload nb64: skip the measured run at the projected budget for eps = 1%%")
    ap.add_argument("--probe", action="store_true",
                    help="no timing: one elementwise copy of 512 MiB (counter calibration) and one batched "
                         "solve, for hardware-counter runs")
    args = ap.parse_args()
    args.nvar_given = args.nvar is not None
    if args.nvar is None:
        args.nvar = 8
    global RESTRICTION
    RESTRICTION = 1 if args.restriction == "spline" else 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU path (cell slabs per GPU) is not "
              "part of it, so --gpus must be 1", file=sys.stderr)
        return 2

    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "slab406":
        return bench_slab(args)
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

    from radchan_amd import VARIANT_BATCHED, config
    from radchan_amd import fields as F
    from radchan_amd import lib as L
    from radchan_amd.tables import synthetic_nb

    def field_of(name):
        q = config(name)
        q.nVar = args.nvar
        if q.grey:
            return q, None, F.channel_field(q, 500.0, 1000.0, noise=0.02, seed=0)
        qtab = tab if (q.nB, q.nT, q.nQ) == (tab.nB, tab.nT, tab.nQ) else synthetic_nb(q.nB, q.nT, q.nQ)
        return q, qtab, F.channel_field(q, 300.0, 1500.0, noise=0.02, seed=0)

    tab = synthetic_nb()
    p, ptab, T = field_of(args.workload)
    lib = L.load()
    bf16 = args.tables == "bf16" and not p.grey
    log(f"bench: workload {args.workload} {p.imax}x{p.jmax}x{p.kmax} photmax {p.photmax} nVar {p.nVar}")

    if args.probe:
        # a device-to-device copy of 512 MiB through the HIP runtime (a blit
        # kernel: reads and writes 512 MiB), then one batched solve
        import ctypes
        s = new_session(p, ptab, [T])
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

    steps = args.steps
    big = p.imax * p.jmax * p.kmax >= 4 << 20
    main_leg = field_legs(args, lib, p, ptab, T, bf16, steps, args.warmup,
                          args.plain_steps if args.plain_steps is not None else (2 if big else 3),
                          fp64_steps=0 if big else 3)
    head = main_leg["head"]

    # the 64^3 field beside the headline, with the measured time to 1%
    second = None
    if not args.no_secondary and args.workload != "nb64" and not p.grey:
        q, qtab, qT = field_of("nb64")
        log("bench: 64x64x64 photmax 200 beside the headline")
        second = field_legs(args, lib, q, qtab, qT, bf16, 10, 2, 3, scaling=False, fp64_steps=3)
        nb = second["head"]
        second["time_to_1pct_measured"] = confirm_leg(lib, q, qtab, qT, second["schedule"], nb["eps"],
                                                      nb["ms_per_step"])
    elif args.workload == "nb64" and not args.no_confirm:
        main_leg["time_to_1pct_measured"] = confirm_leg(lib, p, ptab, T, main_leg["schedule"], head["eps"],
                                                        head["ms_per_step"])

    ck = None if args.no_ckdist else ckdist_leg(not args.no_cpu)

    cpu = None
    cpu_other = None
    if not args.no_cpu:
        try:
            host = host_cpu()
            cpu = cpu_baseline(p, ptab, T, args.cpu_seconds, host)
            log(f"cpu:     {cpu['value'] / 1e6:.2f} Mrays/s on {cpu['cores']} threads ({host['model']}, nproc "
                f"{host['nproc']}), {cpu['one_core']['value'] / 1e6:.3f} Mrays/s on one core")
            cpu_other = cpu_configs(tab, args.cpu_seconds / 3, host)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {e}")

    value = head["rays_per_s"]
    out = {
        "metric": "rays/s (batched solver) and time to 1% std-error of div q_r; batched vs plain, 1 MI355X",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 march, bf16 k-tables, fp64 tallies" if bf16 else "fp32 march, fp32 k-tables, fp64 tallies",
        "data": "synthetic (H2O-like tables, parabolic channel T field with 2% smoothed noise)",
        "config": {"workload": main_leg["workload"], "cells": main_leg["cells"],
                   "rays_per_step": int(head["rays"]), "steps_per_ray": head["steps"] / head["rays"],
                   "schedule": main_leg["schedule"], "graph": not args.no_graph, "restriction": args.restriction},
        "eps": head["eps"],
        "eps_p95_cell": head["eps_p95_cell"],
        "time_to_1pct_s": head["time_to_1pct_s"],
        "batched": main_leg["batched"],
        "plain": main_leg["plain"],
        "speedup_vs_plain": main_leg["speedup_vs_plain"],
        "time_to_1pct_speedup_vs_plain": main_leg["time_to_1pct_speedup_vs_plain"],
        "plain_one_launch_per_band": main_leg["plain_one_launch_per_band"],
        "deterministic_reduction": main_leg["deterministic_reduction"],
        "sobol_directions": main_leg["sobol_directions"],
        "time_to_1pct_speedup_vs_plain_per_band": main_leg["time_to_1pct_speedup_vs_plain_per_band"],
        "eps_scaling": main_leg.get("eps_scaling"),
        "time_to_1pct_measured": main_leg.get("time_to_1pct_measured"),
        "nb64": second,
        "ckdist": ck,
        "roofline": main_leg["roofline"],
        "cpu_baseline": cpu,
        "cpu_baseline_other_configs": cpu_other,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
