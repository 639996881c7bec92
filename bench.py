#!/usr/bin/env python3
"""Benchmark: wet-cell updates per second of the KP step on one MI355X.

Workload (N=1): rain_basin 2048 x 2048 cells, dx = 10 m, diamond-square DEM
0-60 m with a carved channel, n = 0.035, triangular hyetograph 0 -> 80 mm/hr
over 1 h, Green-Ampt K lognormal (median 5 mm/hr, sigma_ln 0.8), psi 0.11,
dtheta 0.3, tile acceleration on, inputs generated with fixed seeds and
resident in HBM before the timed region.

A step is one full Heun time step (both stages, forcing, tile bookkeeping,
dt).  A wet-cell update (WCU) is one interior cell of an active 16 x 14 tile
advanced through one step.  value = WCU over the K timed steps / wall time of
those steps, bracketed by stream synchronisations.

roofline: the dominant kernel of the fused step (the stage-2 march) timed with
HIP events on the handle's stream; bytes per launch = wet cells of the launch
x the per-cell bytes of its contract (KERNEL_BYTES_PER_CELL), against 8 TB/s.
cpu_baseline: the FP32 OpenMP transcription of the same per-cell update
(oracle/liboracle32.so) on the same workload, a bounded number of steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# Bytes per wet cell that each kernel of the fused step must move by its
# contract (rain_basin flags: rain, Green-Ampt, uniform Manning n):
#   march<1>: reads h, hu, hv, BC, BY, BX, K, F (32 B), writes dh, dhu, dhv (12 B)
#   march<2>: reads h, hu, hv, dh, dhu, dhv, BC, BY, BX, K, F (44 B),
#             writes h, hu, hv, F (16 B)
KERNEL_BYTES_PER_CELL = {"march1": 44, "march2": 60}
# algorithmic bytes per WCU of a whole step for this workload (state once per
# stage, one bed field, Green-Ampt F and K; uniform n costs nothing)
STEP_BYTES_PER_WCU = 96


PROFILE = os.path.join(ROOT, "profiles", "fused_step_rain_basin_2048.md")


def measured_traffic(kernel, n_cells):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (calibrated against a copy of known size) for the 2048 x 2048
    workload, as recorded in the committed profile summary; None for other
    workloads or when the summary is absent."""
    if n_cells != 2048 or not os.path.exists(PROFILE):
        return None
    key = f"traffic_bytes_per_launch {kernel}:"
    for line in open(PROFILE):
        if line.startswith(key):
            return float(line[len(key):].strip())
    return None


def build_workload(n_cells):
    from rift_amd import kp, synth
    cfg = synth.rain_basin(n_cells, seed=2)
    flags = kp.TILE_ACCEL | kp.RAIN_AVG
    g = kp.grid_from_rasters(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"],
                             psi=cfg["psi"], dtheta=cfg["dtheta"], flags=flags)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    g.set_series(kp.SER_HYETO, t, v)
    return g, cfg


def cpu_baseline(cfg, g, budget_s=15.0, max_steps=60):
    """Time the FP32 OpenMP transcription on the host cores, continuing from
    the GPU's state (h, hu, hv, F, tile flags, t, dt copied over like a
    hotstart), so both legs time the same step window of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"], psi=cfg["psi"],
                 dtheta=cfg["dtheta"], flags=O.TILE_ACCEL | O.RAIN_AVG, fp32=True)
    o.set_series_hours(O.SER_HYETO, cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    bc = o.field("bc")
    o.set_field("w", g.field("depth") + bc)
    o.set_field("hu", g.field("hu"))
    o.set_field("hv", g.field("hv"))
    o.set_field("F", g.field("F"))
    wet, act, _ = g.tiles()
    o.set_tiles(wet, np.where(act > 0, np.arange(g.ntiles), -np.arange(g.ntiles)))
    t, dt = g.time
    o.set_time(t, dt)
    o.step(1)                       # warm-up (also restores the forcing state)
    cells = 0
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        o.step()
        wet, act, n_act = o.tiles()
        cells += _active_interior_cells(o.nx, o.ny, o.nbx, act)
        steps += 1
    el = time.perf_counter() - t0
    n = cfg["dem"].shape[0] - 1
    return {"value": cells / el, "unit": "WCU/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of rain_basin {n}x{n} continuing from the GPU state after the "
                      f"GPU's timed window (1 untimed step), FP32 OpenMP transcription of the same "
                      f"per-cell update (oracle/liboracle32.so), {el:.1f} s, {threads} threads"}


def ensemble_leg(nscen, seed=4, check_every=64):
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
    """Surrogate screening (rift_amd.surrogate.run) at reduced size: the
    network of the full workload (4 levels, 32 base channels, BF16 on the
    matrix cores) trained on len(train) solver runs at 256 x 256 (dx = 10 m,
    tf = 1 h), then screening len(test) new scenarios with solver fallback
    within 0.1 m of a 0.3 m asset-depth threshold and 5 % held out.  The full
    workload (4000 training runs, 2048 test scenarios) is
    `python -m rift_amd.surrogate`."""
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
                       f"{r['n_test']} test scenarios at 256x256 (reduced from 4000 / 2048), "
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=160)
    ap.add_argument("--cells", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a known-size reference "
                         "for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="scenarios of the batched-ensemble leg (0: skip)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="launch variant of the marching kernels (default: autotuned after the "
                         "warm-up, outside the timed region)")
    ap.add_argument("--surrogate-train", type=int, default=512,
                    help="training scenarios of the surrogate-screening leg (0: skip the leg)")
    ap.add_argument("--surrogate-test", type=int, default=512,
                    help="test scenarios of the surrogate-screening leg")
    ap.add_argument("--surrogate-epochs", type=int, default=12)
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
        print(f"[bench] surrogate scenarios generated ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
              flush=True)

    g, cfg = build_workload(args.cells)
    print(f"[bench] workload ready: {g.nx}x{g.ny} internal, {g.ntiles} tiles", file=sys.stderr,
          flush=True)
    g.step(args.warmup)
    g.sync()
    _, wcu0 = g.wet_cells()
    wet, act, n_act = g.tiles()
    print(f"[bench] warm-up done: {n_act}/{g.ntiles} tiles active", file=sys.stderr, flush=True)
    variant_ms = None
    if args.variant >= 0:
        g.set_variant(args.variant)
        variant = args.variant
    else:
        variant, variant_ms = g.autotune(64)
        print(f"[bench] autotuned launch variant {variant}: "
              + " ".join(f"{m:.4f}" for m in variant_ms) + " ms/step", file=sys.stderr, flush=True)

    # timed region: K steps enqueued back to back, stream synchronised on both sides
    g.sync()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    elapsed = time.perf_counter() - t0
    per_step, wcu1 = g.wet_cells()
    wcu = wcu1 - wcu0

    # kernel durations: a further window with every launch bracketed by HIP
    # events on the handle's stream (kept out of the wall-clock window above)
    kp_steps = min(args.steps, 100)
    g.profile(True)
    _, wk0 = g.wet_cells()
    g.step(kp_steps)
    g.sync()
    _, wk1 = g.wet_cells()
    n1, ms1 = g.kernel_stats("flux1")
    n2, ms2 = g.kernel_stats("stage2_flux")
    ng, msg = g.kernel_stats("grow")
    nt, mst = g.kernel_stats("timestep")
    g.profile(False)
    t_sim, dt = g.time
    mb = g.mass_balance()
    wet, act, n_act = g.tiles()

    value = wcu / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    cells_per_step = (wk1 - wk0) / kp_steps
    kern = {}
    for name, n, ms in (("march1", n1, ms1), ("march2", n2, ms2)):
        avg_s = ms / max(n, 1) * 1e-3
        nbytes = cells_per_step * KERNEL_BYTES_PER_CELL[name]
        kern[name] = {"avg_launch_ms": avg_s * 1e3, "bytes_per_launch": nbytes,
                      "achieved_GBps": nbytes / avg_s / 1e9 if avg_s > 0 else 0.0}
    dom = max(kern, key=lambda k: kern[k]["avg_launch_ms"])
    achieved = kern[dom]["achieved_GBps"] * 1e9
    traffic = measured_traffic("k_march<2>" if dom == "march2" else "k_march<1>", args.cells)
    copy_bw = copy_bandwidth()
    line = {
        "metric": "wet-cell updates/s",
        "value": value,
        "unit": "WCU/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"rain_basin {args.cells}x{args.cells} dx=10m, rain+Green-Ampt+Manning, "
                               "tile acceleration on", "cells": args.cells, "tiles": g.ntiles,
                   "active_tiles": int(n_act), "wet_cells_per_step": wcu / args.steps,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic,
                     "traffic_source": os.path.relpath(PROFILE, ROOT),
                     "kernel": "k_march<2>" if dom == "march2" else "k_march<1>",
                     "bytes_per_launch": kern[dom]["bytes_per_launch"],
                     "avg_launch_ms": kern[dom]["avg_launch_ms"],
                     "traffic_GBps": None if traffic is None
                     else traffic / (kern[dom]["avg_launch_ms"] * 1e-3) / 1e9},
        "hbm_copy": copy_bw,
        "kernels": kern,
        "step_roofline": {"bytes_per_wcu": STEP_BYTES_PER_WCU,
                          "achieved_GBps": value * STEP_BYTES_PER_WCU / 1e9,
                          "frac": value * STEP_BYTES_PER_WCU / HBM_PEAK},
        "kernel_ms_per_step": {"march1": ms1 / kp_steps, "march2": ms2 / kp_steps,
                               "grow": msg / kp_steps, "timestep": mst / kp_steps},
        "launch_variant": variant,
        "variant_ms_per_step": None if variant_ms is None else [float(m) for m in variant_ms],
        "sim_time_s": t_sim,
        "mass_balance_m3": {"stored": mb[0], "infiltrated": mb[1], "rain_in": mb[2]},
    }
    if args.ensemble > 0:
        print("[bench] ensemble ...", file=sys.stderr, flush=True)
        line["ensemble"] = ensemble_leg(args.ensemble)
    else:
        line["ensemble"] = None
    if surr_data is not None:
        print("[bench] surrogate screening ...", file=sys.stderr, flush=True)
        line["surrogate"] = surrogate_leg(surr_data, args.surrogate_epochs)
        del surr_data
    else:
        line["surrogate"] = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(cfg, g)
    else:
        line["cpu_baseline"] = None
    if args.calibrate:
        line["calibration_copy_bytes"] = calibration_copy()
    g.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
