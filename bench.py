#!/usr/bin/env python3
"""Benchmark: wet-cell updates per second of the KP step on one MI355X.

Headline (N=1): coastal surge 8192 x 8192 cells, dx = 5 m, ramp -5 -> +30 m
plus fractal noise (amplitude 4 m), Manning map 0.02 below 0 m / 0.06 above,
37 surge maps every 600 s on the low edge (2.5 sin(pi t / 2 h), clipped at
0) held in device slots, hydrograph 0 -> 500 cfs on a 21 x 21 cell
rectangle, tile acceleration on, tf = 6 h.  The state is spun up to a fixed
simulated time (t = 3 h, 5-25 % wet) outside the timed region and
independent of --warmup; the FP64 volume budget runs over the spin-up and
its residual is reported.  The autotuner then picks the launch variant (on
a copy of the state) and SETTLE_STEPS untimed steps follow (the step runs up
to 9 % slower for ~80 steps after the autotuner's handle is freed; they run
with --variant too, so that every run reaches the same state).  Then
W untimed and K timed steps, inputs resident in HBM, stream synchronised on
both sides.

A step is one full Heun time step (both stages, forcing, surge window, tile
bookkeeping, dt).  A wet-cell update (WCU) is one interior cell of an active
16 x 14 tile advanced through one step.  value = WCU of the K timed steps /
wall time of those steps.

roofline: the dominant kernel of the step (the stage-2 march), timed with HIP
events on the handle's stream in a window right after the timed one; its
algorithmic bytes per launch = wet cells x the stage-2 bytes per WCU of the
workload (STAGE_BYTES: every element once per stage, one corner-DEM bed
field), against 8 TB/s.  step_roofline: the whole step at the workload's
bytes per WCU.  traffic: HBM bytes per launch of that kernel measured with
rocprofv3 FETCH_SIZE / WRITE_SIZE on this same command, read from the
committed profile summary when the run reaches the same state.

Second legs: rain basin 2048 x 2048 at full activity, the batched ensemble
(64 scenarios), surrogate screening; cpu_baseline: the FP32 OpenMP
transcription of the same per-cell update (oracle/liboracle32.so) on a
bounded window of the headline workload.

Profiles: the committed summaries run this command under rocprofv3 with
--graphs 0 (every kernel launched on its own: the same kernels, bit-identical
results).  `python3 bench.py --profile-summary DIR N` summarises rocprofv3's
CSV output over the last N dispatches of each kernel; with N = K + min(K,
100) those are the timed and event windows.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# Algorithmic bytes per WCU (FP32, each element touched once per stage, one
# corner-DEM bed field, U^n kept for stage 2):
#   stage 1: read U^n 12 + write 12 + bed 4 (+ Manning map 4, surge map 4;
#            rain basin: K 4 + F read/write 8)
#   stage 2: read U* 12 + U^n 12 + write 12 + bed 4 (+ Manning 4;
#            rain basin: K 4 + F 12)
STAGE_BYTES = {
    "coastal": {"march1": 36, "march2": 44, "step": 80},
    "rain_basin": {"march1": 40, "march2": 56, "step": 96},
}
SPINUP_S = 3 * 3600.0
SETTLE_STEPS = 128   # untimed steps after the autotuner, before the W warm-up steps
COASTAL_PROFILE = os.path.join(ROOT, "profiles", "fused_step_coastal_8192.md")
RAIN_PROFILE = os.path.join(ROOT, "profiles", "fused_step_rain_basin_2048.md")


def _json_default(o):
    """numpy scalars and arrays in the JSON line."""
    if isinstance(o, np.ndarray):
        return o.tolist()
    if isinstance(o, np.generic):
        return o.item()
    raise TypeError(f"not JSON serialisable: {type(o).__name__}")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def measured_traffic(path, kernel, state, variant):
    """HBM bytes per launch of `kernel` (rocprofv3 FETCH_SIZE / WRITE_SIZE of
    this command, calibrated against a copy of known size) from a committed
    profile summary, when the summary was taken on the same state (same
    `state` key: workload, spin-up, active tiles and start of the timed
    window) with a launch variant of the same traffic class (a line
    `traffic_bytes_per_launch <kernel> variants a,b,..: bytes` listing it);
    else None."""
    if not os.path.exists(path):
        return None
    want = f"state: {state}"
    lines = [ln.rstrip("\n") for ln in open(path)]
    if want not in lines:
        return None
    key = f"traffic_bytes_per_launch {kernel} variants "
    for ln in lines:
        if ln.startswith(key):
            vs, val = ln[len(key):].split(":", 1)
            if variant in {int(v) for v in vs.split(",")}:
                return float(val.strip())
    return None


def build_coastal(n_cells, budget=False):
    """The coastal-surge workload: ramp -5 -> +30 m plus fractal noise, n map
    0.02 / 0.06, 37 surge maps every 600 s held in device slots (generated
    one index at a time, never all on the host), hydrograph 0 -> 500 cfs on
    a 21 x 21 cell rectangle, tile acceleration on.  The series window
    starts at t = 0."""
    from rift_amd import kp, synth
    t0 = time.perf_counter()
    cfg = synth.coastal_surge(n_cells, with_maps=False)
    flags = kp.TILE_ACCEL | kp.SURGE_GRID | kp.HYDROGRAPH | (kp.MASS_BUDGET if budget else 0)
    g = kp.grid_from_rasters(cfg["dem"], cfg["dx"], n=cfg["n"], flags=flags)
    t, v = synth.series_hours_to_seconds(cfg["hydro_hours"], cfg["hydro_cfs"])
    g.set_series(kp.SER_HYDRO, t, v)
    g.set_source_cells(synth.source_cells(g.nx, cfg["source_rect"]))
    nmaps = int(cfg["surge_tf"] // cfg["surge_dt"]) + 1
    g.map_series(kp.MAP_SURGE, nmaps, cfg["surge_dt"], cfg["surge_tf"])
    for k in range(nmaps):
        g.load_map(kp.MAP_SURGE, k, kp.map_cells(synth.surge_map(n_cells, k, cfg["surge_dt"])))
    g.start_maps(kp.MAP_SURGE, 0.0)
    cfg["nmaps"] = nmaps
    log(f"coastal {n_cells}x{n_cells} ready: {g.ntiles} tiles, {nmaps} surge maps on the device "
        f"({time.perf_counter() - t0:.0f} s)")
    return g, cfg


def build_rain_basin(n_cells):
    from rift_amd import kp, synth
    cfg = synth.rain_basin(n_cells, seed=2)
    flags = kp.TILE_ACCEL | kp.RAIN_AVG
    g = kp.grid_from_rasters(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"],
                             psi=cfg["psi"], dtheta=cfg["dtheta"], flags=flags)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    g.set_series(kp.SER_HYETO, t, v)
    return g, cfg


def spin_up(g, t_target, chunk=4096, label="spin-up", piece=0):
    """Steps until the simulated time reaches t_target (whole chunks of
    graph-replayed steps, then single steps); returns the step count.  With
    piece > 0 each chunk runs as pieces of at most `piece` steps with a
    stream synchronisation between them (the same steps and states; it
    bounds the dispatches in flight, for rocprofv3 counter collection)."""
    steps = 0
    t = g.time[0]
    w0 = time.perf_counter()
    while t < t_target:
        _, dt = g.time
        n = max(1, min(chunk, int(0.5 * (t_target - t) / max(dt, 1e-9))))
        if piece > 0:
            for k in range(0, n, piece):
                g.step(min(piece, n - k))
                g.sync()
        else:
            g.step(n)
        steps += n
        t = g.time[0]
        if n == chunk:
            _, _, nact = g.tiles()
            log(f"{label}: t = {t:.0f} s, {steps} steps, {nact} active tiles "
                f"({time.perf_counter() - w0:.0f} s)")
    return steps


def timed_windows(g, steps, warmup, variant, stage_bytes):
    """W untimed steps, then K timed steps (wall clock between stream
    synchronisations), then a window of min(K, 100) steps with every kernel
    bracketed by HIP events on the handle's stream."""
    if variant >= 0:
        g.set_variant(variant)
        variant_ms = None
    else:
        variant, variant_ms = g.autotune(64)
        log("autotuned launch variant %d: %s ms/step" % (variant, " ".join(f"{m:.4f}" for m in variant_ms)))
    # after the autotuner (a second handle of the same size created, stepped
    # and freed) the step runs up to 9 % slower for ~80 steps; the settle
    # steps run with a given variant too, so that both reach the same state
    g.step(SETTLE_STEPS)
    g.step(warmup)
    g.sync()
    t_start, _ = g.time
    _, wcu0 = g.wet_cells()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    elapsed = time.perf_counter() - t0
    _, wcu1 = g.wet_cells()
    kp_steps = min(steps, 100)
    g.profile(True)
    _, wk0 = g.wet_cells()
    g.step(kp_steps)
    g.sync()
    _, wk1 = g.wet_cells()
    stats = {c: g.kernel_stats(c) for c in ("flux1", "stage2_flux", "grow", "timestep")}
    g.profile(False)
    wet, act, n_act = g.tiles()
    wcu = wcu1 - wcu0
    value = wcu / elapsed
    cells_per_launch = (wk1 - wk0) / kp_steps
    kern = {}
    for name, cls in (("march1", "flux1"), ("march2", "stage2_flux")):
        n, ms = stats[cls]
        avg_s = ms / max(n, 1) * 1e-3
        nbytes = cells_per_launch * stage_bytes[name]
        kern[name] = {"avg_launch_ms": avg_s * 1e3, "launches": n, "bytes_per_launch": nbytes,
                      "achieved_GBps": nbytes / avg_s / 1e9 if avg_s > 0 else 0.0}
    dom = max(kern, key=lambda k: kern[k]["avg_launch_ms"])
    return {
        "value": value, "elapsed": elapsed, "ms_per_step": 1e3 * elapsed / max(steps, 1),
        "wcu": wcu, "t_start": t_start, "n_act": int(n_act), "kern": kern, "dom": dom,
        "cells_per_launch": cells_per_launch,
        "kernel_ms_per_step": {"march1": stats["flux1"][1] / kp_steps,
                               "march2": stats["stage2_flux"][1] / kp_steps,
                               "grow_track": stats["grow"][1] / kp_steps,
                               "timestep": stats["timestep"][1] / kp_steps},
        "variant": variant, "variant_ms": variant_ms,
    }


def roofline(w, stage_bytes, traffic):
    k = w["kern"][w["dom"]]
    achieved = k["achieved_GBps"] * 1e9
    name = "k_march<2>" if w["dom"] == "march2" else "k_march<1>"
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": name,
            "bytes_per_wcu": stage_bytes[w["dom"]], "bytes_per_launch": k["bytes_per_launch"],
            "avg_launch_ms": k["avg_launch_ms"],
            "traffic_over_algorithmic": None if traffic is None else traffic / k["bytes_per_launch"]}


def coastal_cpu_baseline(cfg, g, budget_s=20.0, max_steps=60):
    """The FP32 OpenMP transcription (oracle/liboracle32.so) continuing from
    the GPU's state (w, hu, hv, tile flags, t, dt, as a hotstart; the surge
    series restarts its window at t with the maps around it) on the host
    cores: a bounded window of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from rift_amd import kp, synth
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n_cells = cfg["dem"].shape[0] - 1
    t_setup = time.perf_counter()
    o = O.Oracle(cfg["dem"], cfg["dx"], n=cfg["n"],
                 flags=O.TILE_ACCEL | O.HYDROGRAPH | O.SURGE_GRID, fp32=True)
    o.set_sources(synth.source_cells(o.nx, cfg["source_rect"]))
    o.set_series_hours(O.SER_HYDRO, cfg["hydro_hours"], cfg["hydro_cfs"])
    for name in ("w", "hu", "hv"):
        o.set_field(name, g.field(name))
    wet, act, _ = g.tiles()
    o.set_tiles(wet, np.where(act > 0, np.arange(g.ntiles), -np.arange(g.ntiles)))
    t, dt = g.time
    o.set_time(t, dt)
    k0 = g.map_state(kp.MAP_SURGE)[0]
    maps = np.stack([synth.surge_map(n_cells, k, cfg["surge_dt"]) for k in range(k0, k0 + 3)])
    o.set_map_series(O.MAP_SURGE, maps, cfg["surge_dt"], cfg["surge_tf"], first=k0)
    del maps
    o.step(1)                       # warm-up (also restores the forcing state)
    setup = time.perf_counter() - t_setup
    cells = 0
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        o.step()
        _, oact, _ = o.tiles()
        cells += _active_interior_cells(o.nx, o.ny, o.nbx, oact)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": cells / el, "unit": "WCU/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{steps} steps of coastal {n_cells}x{n_cells} continuing from the GPU state at "
                      f"t = {t:.0f} s (after the GPU's timed windows; 1 untimed step), FP32 OpenMP "
                      f"transcription of the same per-cell update and tile logic "
                      f"(oracle/liboracle32.so), {el:.1f} s, {threads} threads ({setup:.0f} s set-up)"}


def ensemble_leg(nscen, seed=4, check_every=64, graphs=True):
    """The batched-ensemble workload (rift_amd.synth.ensemble_scenarios: sides
    512..2048, dx = 10 m, tf = 1 h, per-scenario rain peak, Manning n and
    Green-Ampt K) advanced to every scenario's end time by rift_ensemble_run;
    FP64 volume budgets per scenario.  Scenarios/hour = scenarios / wall time
    of the run (inputs resident in HBM) x 3600."""
    from rift_amd import kp, synth
    t_gen = time.perf_counter()
    scen = synth.ensemble_scenarios(nscen, seed)
    flags = kp.TILE_ACCEL | kp.RAIN_AVG | kp.MASS_BUDGET
    spec = [dict(dem=m["dem"], dx=m["dx"], n_const=m["n_const"], K=m["K"], psi=m["psi"],
                 dtheta=m["dtheta"], flags=flags) for m in scen]
    e = kp.Ensemble(spec)
    for i, (g, m) in enumerate(zip(e.members, scen)):
        t, v = synth.series_hours_to_seconds(m["hours"], m["mmhr"])
        g.set_series(kp.SER_HYETO, t, v)
        e.set_tf(i, m["tf"])
        if not graphs:
            g.set_graphs(False)
    cells = sum(m["side"] ** 2 for m in scen)
    del scen, spec
    setup_s = time.perf_counter() - t_gen
    print(f"[bench] ensemble of {nscen} ready ({cells / 1e6:.1f} M cells, {setup_s:.0f} s set-up)",
          file=sys.stderr, flush=True)
    e.members[0].sync()
    t0 = time.perf_counter()
    steps = e.run(check_every=check_every)
    e.members[0].sync()
    wall = time.perf_counter() - t0
    wcu = sum(g.wet_cells()[1] for g in e.members)
    worst = 0.0
    for g in e.members:
        b = g.budget()
        worst = max(worst, abs(b[8]) / max(np.abs(b[:8]).max(), 1e-300))
    out = {"scenarios": nscen, "cells": cells, "ensemble_steps": steps, "wall_s": wall,
           "scenarios_per_hour": nscen / wall * 3600.0, "wcu_per_s": wcu / wall,
           "max_budget_residual_rel": worst, "setup_s": setup_s,
           "workload": "sides {512..2048} dx=10m, tf=1h, lognormal rain peak (median 50 mm/hr), "
                       "n~U(0.02,0.08), Green-Ampt K lognormal; plain batched solver"}
    e.close()
    return out


def surrogate_leg(data, epochs):
    """Surrogate screening (rift_amd.surrogate.run): the U-Net (4 levels, 32
    base channels, BF16 on the matrix cores) trained on len(train) solver
    runs at 256 x 256 (dx = 10 m, tf = 1 h; 4000 by default, the full
    workload), then screening len(test) new scenarios (2048) with solver
    fallback within 0.1 m of a 0.3 m asset-depth threshold and 5 % held out."""
    from rift_amd import surrogate
    r = surrogate.run(epochs=epochs, data=data,
                      log=lambda m: print("[bench] surrogate: " + m, file=sys.stderr, flush=True))
    keep = ["n_train", "n_test", "epochs", "train_solver_s", "train_s", "train_tflops", "infer_s",
            "infer_tflops", "fallback", "heldout", "solver_runs", "screened_per_hour", "plain_per_hour",
            "gain", "gain_without_heldout_est", "heldout_errors", "all_errors_surrogate",
            "all_errors_screened", "decision_accuracy", "flooded_frac"]
    out = {k: r.get(k) for k in keep}
    out["final_train_loss"] = r["train_loss"][-1] if r["train_loss"] else None
    st = surrogate.step_timing(16)
    out["step_timing"] = st
    out["roofline"] = {"bound": "mfma", "kernel": "training step (k_gemm, k_wgrad, ...)",
                       "achieved": st["train_tflops"], "peak": surrogate.BF16_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": st["train_tflops"] / surrogate.BF16_PEAK_TFLOPS}
    out["workload"] = (f"U-Net 4 levels x 32 base ch, BF16 MFMA; {r['n_train']} training runs / "
                       f"{r['n_test']} test scenarios at 256x256, "
                       "threshold 0.3 m +- 0.1 m at the central 8x8 cells, 5% held out")
    return out


def _active_interior_cells(nx, ny, nbx, act):
    """Interior cells of the tiles in an active list (reference encoding)."""
    total = 0
    for t in act:
        if t < 0:
            continue
        bx, by = t % nbx, t // nbx
        i0, j0 = bx * 16 + 2, by * 14 + 2
        ni = min(i0 + 16, nx - 2) - i0
        nj = min(j0 + 14, ny - 2) - j0
        if ni > 0 and nj > 0:
            total += ni * nj
    return total


def calibration_copy(nbytes=256 << 20):
    """One hipMemcpyDtoD of nbytes (runs as the runtime's copy kernel)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    return nbytes


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """What a plain device-to-device copy achieves on this GPU: `reps`
    hipMemcpyAsync of nbytes between two fresh allocations on the null
    stream, timed with HIP events on that stream after two untimed copies;
    bytes moved = 2 x nbytes per copy (read + write).  It puts the marching
    kernels' measured HBM traffic beside an attainable rate, not the
    published peak alone."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    assert hip.hipMalloc(ctypes.byref(a), n) == 0
    assert hip.hipMalloc(ctypes.byref(b), n) == 0
    try:
        assert hip.hipMemset(a, 1, n) == 0
        assert hip.hipEventCreate(ctypes.byref(e0)) == 0
        assert hip.hipEventCreate(ctypes.byref(e1)) == 0
        for _ in range(2):
            assert hip.hipMemcpyAsync(b, a, n, 3, None) == 0      # hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            assert hip.hipMemcpyAsync(b, a, n, 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return {"GBps": 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9, "bytes_per_copy": nbytes,
            "copies": reps, "ms_per_copy": ms.value / reps}


def coastal_leg(args):
    from rift_amd import kp
    g, cfg = build_coastal(args.cells, budget=bool(args.spinup_budget))
    if not args.graphs:
        g.set_graphs(False)
    spin_steps = spin_up(g, SPINUP_S, label="coastal spin-up", piece=0 if args.graphs else 32)
    b = g.budget() if args.spinup_budget else [float("nan")] * 9
    g.set_budget(False)
    t_spin, _ = g.time
    h = g.interior(g.field("depth"))
    wet_frac = float((h > 1.19209e-07).mean())
    _, _, n_act0 = g.tiles()
    log(f"coastal spun up: t = {t_spin:.1f} s after {spin_steps} steps, {n_act0}/{g.ntiles} tiles active, "
        f"{wet_frac:.3f} of cells wet")
    log("budget m^3: stored %.6e infiltrated %.3e at reset %.6e rain %.3e source %.6e drain %.3e "
        "outflow %.6e exchange %.6e residual %.3e" % tuple(b))
    sb = STAGE_BYTES["coastal"]
    w = timed_windows(g, args.steps, args.warmup, args.variant, sb)
    state = (f"coastal {args.cells} t0={SPINUP_S:.0f} spin={spin_steps} active={n_act0} "
             f"settle={SETTLE_STEPS} warmup={args.warmup} t_timed={w['t_start']:.4f}")
    traffic = measured_traffic(COASTAL_PROFILE, "k_march<2>" if w["dom"] == "march2" else "k_march<1>",
                               state, w["variant"])
    k0, k1, f, miss = g.map_state(kp.MAP_SURGE)
    scale = max(abs(b[0]) + abs(b[1]), abs(b[2]), abs(b[4]), abs(b[6]), 1e-300)
    out = {
        "w": w, "g": g, "cfg": cfg, "state": state,
        "config": {"workload": f"coastal_surge {args.cells}x{args.cells} dx=5m: ramp -5..+30 m + fractal "
                               "noise, Manning map 0.02/0.06, surge maps every 600 s (device slots), "
                               "hydrograph 0-500 cfs, tile acceleration on",
                   "cells": args.cells, "tiles": g.ntiles, "spinup_to_s": SPINUP_S,
                   "spinup_steps": spin_steps, "t_timed_start_s": w["t_start"],
                   "active_tiles": w["n_act"], "active_tile_frac": w["n_act"] / g.ntiles,
                   "wet_cell_frac_at_spinup": wet_frac,
                   "wet_cells_per_step": w["wcu"] / max(args.steps, 1),
                   "surge_window": [k0, k1, f], "surge_maps_missing": miss,
                   "parallelism": "single GPU"},
        "roofline": roofline(w, sb, traffic),
        "step_roofline": {"bytes_per_wcu": sb["step"], "achieved_GBps": w["value"] * sb["step"] / 1e9,
                          "frac": w["value"] * sb["step"] / HBM_PEAK},
        "mass_budget": None if not args.spinup_budget else
        {"window": f"t = 0 .. {t_spin:.0f} s ({spin_steps} steps), FP64 on the device",
         "stored_m3": b[0], "source_in_m3": b[4], "outflow_m3": b[6],
         "surge_and_clamp_exchange_m3": b[7], "residual_m3": b[8],
         "residual_rel": abs(b[8]) / scale},
    }
    return out


def rain_basin_leg(args):
    """Rain basin 2048 x 2048 with every tile active (the rain keeps every
    cell wet, so the active set fills from tile 0 within a few hundred
    steps): K timed steps at full activity."""
    g, cfg = build_rain_basin(2048)
    if not args.graphs:
        g.set_graphs(False)
    steps = 0
    while True:
        g.step(50)
        steps += 50
        _, _, n = g.tiles()
        if n == g.ntiles or steps >= 1000:
            break
    log(f"rain basin at full activity after {steps} steps ({n}/{g.ntiles} tiles)")
    sb = STAGE_BYTES["rain_basin"]
    w = timed_windows(g, max(args.steps, 50), 5, args.rain_variant, sb)
    g.close()
    state = (f"rain_basin 2048 spin={steps} active={n} settle={SETTLE_STEPS} warmup=5 "
             f"t_timed={w['t_start']:.4f}")
    traffic = measured_traffic(RAIN_PROFILE, "k_march<2>" if w["dom"] == "march2" else "k_march<1>",
                               state, w["variant"])
    return {"value": w["value"], "unit": "WCU/s", "ms_per_step": w["ms_per_step"],
            "workload": "rain_basin 2048x2048 dx=10m, rain + Green-Ampt + Manning, every tile active",
            "spinup_steps": steps, "active_tiles": w["n_act"], "profile_state": state,
            "roofline": roofline(w, sb, traffic),
            "step_roofline": {"bytes_per_wcu": sb["step"], "frac": w["value"] * sb["step"] / HBM_PEAK},
            "kernel_ms_per_step": w["kernel_ms_per_step"], "launch_variant": w["variant"],
            "variant_ms_per_step": w["variant_ms"]}


def _profile_rows(root):
    """Per kernel, the dispatches of rocprofv3 CSV output under `root`, in
    dispatch order: kernel trace -> [duration ns], counter collection ->
    {counter: [value]}."""
    import csv
    import glob
    durs, ctrs = {}, {}
    for path in sorted(glob.glob(os.path.join(root, "**", "*.csv"), recursive=True)):
        base = os.path.basename(path)
        if base.endswith("kernel_trace.csv"):
            with open(path) as fh:
                rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Dispatch_Id"]))
            for r in rows:
                durs.setdefault(r["Kernel_Name"], []).append(
                    int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        elif base.endswith("counter_collection.csv"):
            with open(path) as fh:
                rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Dispatch_Id"]))
            for r in rows:
                d = ctrs.setdefault(r["Kernel_Name"], {})
                d.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return durs, ctrs


def profile_summary(root, last, calib_bytes=256 << 20):
    """Markdown tables from rocprofv3 output under `root`: per kernel the
    average duration over all dispatches and over the last `last` (the timed
    and event windows of a bench run; earlier dispatches are the spin-up and
    the autotuner), and per counter the average of the last `last`.  With
    FETCH_SIZE / WRITE_SIZE present, the run's last __amd_rocclr_copyBuffer
    dispatch (the --calibrate copy of calib_bytes) scales KiB to bytes."""
    durs, ctrs = _profile_rows(root)
    out = []
    if durs:
        out += ["| kernel | calls | avg us (all) | avg us (last %d) |" % last, "|---|---|---|---|"]
        for name, d in sorted(durs.items(), key=lambda kv: -sum(kv[1][-last:])):
            tail = d[-last:]
            out.append(f"| `{name}` | {len(d)} | {sum(d) / len(d) / 1e3:.2f} | "
                       f"{sum(tail) / len(tail) / 1e3:.2f} |")
    scale = {}
    copy = [k for k in ctrs if "copyBuffer" in k]
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = [v for k in copy for v in ctrs[k].get(c, [])[-1:]]
        if vals:
            scale[c] = calib_bytes / (vals[0] * 1024.0)
            out.append(f"\ncalibration {c}: last copyBuffer dispatch {vals[0]:.0f} KiB for "
                       f"{calib_bytes} B: factor {scale[c]:.4f}")
    if ctrs:
        names = sorted({c for d in ctrs.values() for c in d})
        out += ["", "| kernel | " + " | ".join(f"{c} (avg of last {last})" for c in names) + " |",
                "|---|" + "---|" * len(names)]
        for name, d in sorted(ctrs.items()):
            if "copyBuffer" in name:
                continue
            cells = []
            for c in names:
                v = d.get(c, [])[-last:]
                cells.append(f"{sum(v) / len(v):.0f}" if v else "")
            out.append(f"| `{name}` | " + " | ".join(cells) + " |")
        for name, d in sorted(ctrs.items()):
            if "copyBuffer" in name or not scale:
                continue
            nb = 0.0
            for c, f in scale.items():
                v = d.get(c, [])[-last:]
                if v:
                    nb += sum(v) / len(v) * 1024.0 * f
            short = name.split("(")[0].replace("void ", "").replace("rift::", "")
            out.append(f"\nhbm_bytes_per_launch {short}: {nb:.0f}")
    return "\n".join(out)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--profile-summary":
        # python3 bench.py --profile-summary DIR LAST: summarise rocprofv3 output (no GPU)
        print(profile_summary(sys.argv[2], int(sys.argv[3])))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 256 MiB device-to-device copy in a profiled region (a "
                         "known-size reference for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--rain-basin", type=int, default=1, help="run the 2048^2 full-activity leg (0: skip)")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="scenarios of the batched-ensemble leg (0: skip)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="1: steps replayed from captured HIP graphs (default); 0: every kernel "
                         "launched on its own (same kernels and results)")
    ap.add_argument("--spinup-budget", type=int, default=1,
                    help="1: the FP64 volume budget runs over the spin-up and its residual is "
                         "reported (default); 0: off (the same states; rocprofv3 --pmc passes use it)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="launch variant of the marching kernels (default: autotuned after the "
                         "spin-up, outside the timed region)")
    ap.add_argument("--rain-variant", type=int, default=-1,
                    help="launch variant of the rain-basin leg (default: autotuned)")
    ap.add_argument("--surrogate-train", type=int, default=4000,
                    help="training scenarios of the surrogate-screening leg (0: skip the leg)")
    ap.add_argument("--surrogate-test", type=int, default=2048,
                    help="test scenarios of the surrogate-screening leg")
    ap.add_argument("--surrogate-epochs", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the solver runs one domain per GPU)",
              file=sys.stderr)
        sys.exit(2)

    surr_data = None
    if args.surrogate_train > 0:
        # scenario inputs come from a process pool, before this process
        # touches the GPU
        from rift_amd import surrogate
        t0 = time.perf_counter()
        surr_data = surrogate.prepare(args.surrogate_train, args.surrogate_test, workers=8)
        log(f"surrogate scenarios generated ({time.perf_counter() - t0:.0f} s)")

    c = coastal_leg(args)
    w, g = c["w"], c["g"]
    line = {
        "metric": "wet-cell updates/s",
        "value": w["value"],
        "unit": "WCU/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": c["config"],
        "roofline": c["roofline"],
        "step_roofline": c["step_roofline"],
        "kernels": w["kern"],
        "kernel_ms_per_step": w["kernel_ms_per_step"],
        "launch_variant": w["variant"],
        "variant_ms_per_step": None if w["variant_ms"] is None else [float(m) for m in w["variant_ms"]],
        "variant_spread": None if w["variant_ms"] is None
        else float(max(w["variant_ms"]) / min(w["variant_ms"]) - 1.0),
        "mass_budget": c["mass_budget"],
        "profile_state": c["state"],
    }
    line["hbm_copy"] = copy_bandwidth()
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = coastal_cpu_baseline(c["cfg"], g)
    else:
        line["cpu_baseline"] = None
    g.close()
    del c, g
    if args.rain_basin:
        log("rain basin 2048 at full activity ...")
        line["rain_basin_2048"] = rain_basin_leg(args)
    if args.ensemble > 0:
        log("ensemble ...")
        line["ensemble"] = ensemble_leg(args.ensemble, graphs=bool(args.graphs))
    else:
        line["ensemble"] = None
    if surr_data is not None:
        log("surrogate screening ...")
        line["surrogate"] = surrogate_leg(surr_data, args.surrogate_epochs)
        del surr_data
    else:
        line["surrogate"] = None
    if args.calibrate:
        line["calibration_copy_bytes"] = calibration_copy()
    print(json.dumps(line, default=_json_default))


if __name__ == "__main__":
    main()
