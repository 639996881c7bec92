"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Tolerances (FP32 device state vs the FP64 oracle of the reference scheme):
  - face fluxes / dU of one flux pass: |diff| <= 2e-5 * max|dU| + 1e-7
  - dt of every step: relative 2e-5 on smooth flows; 5e-3 for thin sheet flow
    on rough terrain, where the face depth h + (BC - B_face) + slope is a
    small difference of ~0.5 m terms and FP32 resolves it to ~1e-7 m (the
    FP32 CPU transcription of the reference deviates by the same amount)
  - depth after a whole run: max |dh| <= 1e-4 m (the target bound)
  - wet / active tile sets: identical at every step (integer work)
"""
import numpy as np
import pytest

import oracle_lib as O
from rift_amd import synth

pytestmark = pytest.mark.gpu


def _pair(rift, dem, dx, *, h0=None, n=None, n_const=0.0, K=None, psi=0.0, dtheta=0.0,
          courant=0.25, oflags=O.TILE_ACCEL, mode=O.FAITHFUL, t0=0.0, per_phase=False):
    gflags = rift.TILE_ACCEL if (oflags & O.TILE_ACCEL) else 0
    for ob, gb in ((O.RAIN_AVG, rift.RAIN_AVG), (O.EULER, rift.EULER), (O.SAVE_MAX, rift.SAVE_MAX),
                   (O.SAVE_ARRIVAL, rift.SAVE_ARRIVAL), (O.HYDROGRAPH, rift.HYDROGRAPH),
                   (O.DRAIN_AVG, rift.DRAIN_AVG)):
        if oflags & ob:
            gflags |= gb
    if mode == O.TEXTBOOK:
        gflags |= rift.TEXTBOOK
    o = O.Oracle(dem, dx, h0=h0, n=n, n_const=n_const, K=K, psi=psi, dtheta=dtheta,
                 courant=courant, flags=oflags, mode=mode, t0=t0)
    g = rift.grid_from_rasters(dem, dx, h0=h0, n=n, n_const=n_const, K=K, psi=psi,
                               dtheta=dtheta, courant=courant, flags=gflags, t0=t0,
                               per_phase=per_phase)
    return o, g


def _active_oracle(o):
    wet, act, n = o.tiles()
    return {int(v) for v in act if v > -1}


def test_first_flux_dam_break(rift):
    cfg = synth.dam_break(256)
    o, g = _pair(rift, cfg["dem"], 1.0, h0=cfg["h0"])
    np.testing.assert_allclose(g.interior(g.field("depth")), o.interior(o.depth()), atol=1e-7)
    for ph in ("grow", "compute_fluxes"):
        o.phase(ph)
        g.phase(ph)
    assert g.active_set() == _active_oracle(o)
    for f in ("dw", "dhu", "dhv"):
        a, b = g.interior(g.field(f)), o.interior(o.field(f))
        tol = 2e-5 * max(np.abs(b).max(), 1e-3) + 1e-7
        assert np.abs(a - b).max() <= tol, (f, np.abs(a - b).max(), tol)
    gmx, gmy = g.tile_speeds()
    omx, omy = o.tile_speeds()
    np.testing.assert_allclose(gmx, omx, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(gmy, omy, rtol=2e-6, atol=1e-7)
    g.close()


def _run_compare(o, g, nsteps=None, t_end=None, check_tiles=True, dt_rtol=2e-5):
    k = 0
    max_dt_err = 0.0
    while True:
        if nsteps is not None and k >= nsteps:
            break
        if t_end is not None and o.time[0] >= t_end:
            break
        o.step()
        g.step()
        to, dto = o.time
        tg, dtg = g.time
        max_dt_err = max(max_dt_err, abs(dtg - dto) / dto)
        assert abs(dtg - dto) <= dt_rtol * dto, (k, dto, dtg)
        if check_tiles:
            assert g.active_set() == _active_oracle(o), k
            gw = g.tiles()[0]
            ow = o.tiles()[0]
            np.testing.assert_array_equal(gw, ow)
        k += 1
    return k, max_dt_err


PATHS = pytest.mark.parametrize("per_phase", [False, True], ids=["fused", "per_phase"])


@PATHS
def test_dam_break_full_run(rift, per_phase):
    """256 x 256 dam break to t = 20 s (about 340 steps)."""
    cfg = synth.dam_break(256)
    o, g = _pair(rift, cfg["dem"], 1.0, h0=cfg["h0"], per_phase=per_phase)
    steps, dterr = _run_compare(o, g, t_end=20.0)
    assert steps > 300
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    for f in ("hu", "hv"):
        d = np.abs(g.interior(g.field(f)) - o.interior(o.field(f))).max()
        assert d <= 2e-4, (f, d)
    g.close()


@PATHS
def test_dam_break_textbook_mode(rift, per_phase):
    cfg = synth.dam_break(128)
    o, g = _pair(rift, cfg["dem"], 1.0, h0=cfg["h0"], mode=O.TEXTBOOK, per_phase=per_phase)
    _run_compare(o, g, nsteps=120)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    g.close()


@PATHS
def test_rain_basin_infiltration_friction(rift, per_phase):
    """The rain_basin recipe at 128 x 128: rain, Green-Ampt, Manning, dry start."""
    cfg = synth.rain_basin(128, seed=2)
    o, g = _pair(rift, cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"],
                 psi=cfg["psi"], dtheta=cfg["dtheta"], oflags=O.TILE_ACCEL | O.RAIN_AVG,
                 per_phase=per_phase)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    o.set_series_hours(O.SER_HYETO, cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    g.set_series(rift.SER_HYETO, t, v)
    _run_compare(o, g, nsteps=150, dt_rtol=5e-3)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    dF = np.abs(g.interior(g.field("F")) - o.interior(o.field("F"))).max()
    assert dF <= 1e-4, dF
    g.close()


def test_euler_mode_and_diagnostics(rift):
    cfg = synth.dam_break(64)
    flags = O.TILE_ACCEL | O.EULER | O.SAVE_MAX | O.SAVE_ARRIVAL
    o, g = _pair(rift, cfg["dem"], 1.0, h0=cfg["h0"], n_const=0.02, oflags=flags)
    _run_compare(o, g, nsteps=60)
    for f, tol in (("depth", 1e-4), ("h_max", 1e-4), ("q_max", 2e-4)):
        a = g.interior(g.field(f))
        b = o.interior(o.depth() if f == "depth" else o.field(f))
        assert np.abs(a - b).max() <= tol, f
    g.close()


@PATHS
def test_hydrograph_source_rectangle(rift, per_phase):
    N = 64
    dem = np.zeros((N + 1, N + 1))
    nx = N + 4          # (N+1) DEM points + 3
    idx = [j * nx + i for i in range(30, 34) for j in range(20, 23)]
    hours = np.array([0.0, 0.01, 0.05], dtype=np.float32)
    cfs = np.array([0.0, 50.0, 50.0], dtype=np.float32)
    o, g = _pair(rift, dem, 2.0, n_const=0.03, oflags=O.TILE_ACCEL | O.HYDROGRAPH,
                 per_phase=per_phase)
    o.set_sources(idx)
    o.set_series_hours(O.SER_HYDRO, hours, cfs)
    g.set_source_cells(idx)
    t, v = synth.series_hours_to_seconds(hours, cfs)
    g.set_series(rift.SER_HYDRO, t, v)
    _run_compare(o, g, nsteps=80)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    g.close()


@pytest.mark.parametrize("which", ["hyeto", "drain", "hydro"])
def test_series_rates_bit_exact(rift, which):
    """Source::InterpolateRate + the unit division, evaluated on the device at
    the same times as the oracle, including the stale re-division outside the
    table and the appended zero point (bit-exact float)."""
    hours = np.array([0.25, 0.5, 0.9, 1.3], dtype=np.float32)
    vals = np.array([3.0, 80.0, 17.5, 4.0], dtype=np.float32)
    oflag = {"hyeto": O.RAIN_AVG, "drain": O.DRAIN_AVG, "hydro": O.HYDROGRAPH}[which]
    sid = {"hyeto": 0, "drain": 1, "hydro": 2}[which]
    col = {"hyeto": 0, "drain": 1, "hydro": 2}[which]
    dem = np.zeros((17, 17))
    o, g = _pair(rift, dem, 3.0, oflags=O.TILE_ACCEL | oflag)
    idx = [8 * 20 + 8, 8 * 20 + 9]
    if which == "hydro":
        o.set_sources(idx)
        g.set_source_cells(idx)
    o.set_series_hours(sid, hours, vals)
    t, v = synth.series_hours_to_seconds(hours, vals)
    g.set_series(sid, t, v)
    times = [0.0, 100.0, 900.0, 901.5, 1799.99, 2500.0, 3240.0, 4000.7, 4680.0, 5100.0,
             5500.0, 9000.0, 20000.0, 20001.0]
    for tt in times:
        o.set_time(tt, 0.5)
        g.set_time(tt, 0.5)
        o.phase("update_source")
        g.phase("update_source")
        go, gg = o.rates(), g.rates()
        if which == "hydro":
            # per-cell source rate: Q/dx/dy/N in double, stored as float
            o_cell = np.float32(np.float64(go[2]) / 3.0 / 3.0 / len(idx))
            assert gg[2] == o_cell, (tt, gg[2], o_cell)
        else:
            assert gg[col] == go[col], (tt, gg[col], go[col])
    g.close()


def test_fused_and_per_phase_paths_agree(rift):
    """Both step paths through the C ABI on a wet/dry front with a rough bed
    and a Manning map: same tile sets every step, states within FP32
    rounding of each other."""
    cfg = synth.rain_basin(96, seed=7)
    dem = cfg["dem"]
    h0 = np.maximum(35.0 - dem, 0.0)
    n = 0.02 + 0.05 * synth.uniform(8, dem.size).reshape(dem.shape)
    a = rift.grid_from_rasters(dem, 10.0, h0=h0, n=n, flags=rift.TILE_ACCEL)
    b = rift.grid_from_rasters(dem, 10.0, h0=h0, n=n, flags=rift.TILE_ACCEL, per_phase=True)
    for k in range(60):
        a.step()
        b.step()
        assert a.active_set() == b.active_set(), k
        np.testing.assert_array_equal(a.tiles()[0], b.tiles()[0])
        ta, dta = a.time
        tb, dtb = b.time
        assert abs(dta - dtb) <= 1e-5 * dtb, (k, dta, dtb)
    dh = np.abs(a.interior(a.field("depth")) - b.interior(b.field("depth"))).max()
    assert dh <= 2e-5, dh
    a.close()
    b.close()


def test_graph_replays_match_eager_launches(rift):
    """The fused step replayed from captured graphs (16-step and 1-step
    graphs) equals the same step launched kernel by kernel, bit for bit, on
    every replay, with downloads (eager work and synchronisation) between."""
    cfg = synth.rain_basin(96, seed=3)
    kw = dict(n_const=cfg["n_const"], K=cfg["K"], psi=cfg["psi"], dtheta=cfg["dtheta"],
              flags=rift.TILE_ACCEL | rift.RAIN_AVG)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    a = rift.grid_from_rasters(cfg["dem"], cfg["dx"], **kw)
    b = rift.grid_from_rasters(cfg["dem"], cfg["dx"], **kw)
    for g in (a, b):
        g.set_series(rift.SER_HYETO, t, v)
    b.set_graphs(False)
    for rep, n in enumerate((16, 3, 16, 1, 19)):
        a.step(n)
        b.step(n)
        assert a.time == b.time, rep
        for f in ("depth", "hu", "hv", "F"):
            np.testing.assert_array_equal(a.field(f), b.field(f))
        np.testing.assert_array_equal(a.tiles()[0], b.tiles()[0])
    a.close()
    b.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_dt_in_stage1_march_matches_dt_kernel(rift, graphs):
    """dt formed by the last workgroup of the stage-1 march (default) equals
    the dt kernel's (RIFT_OPT_DT_KERNEL 1) bit for bit: time and dt every
    replay, state, tile sets, wet-cell counts and the next steps' forcing
    (hyetograph, hydrograph), the FP64 budget terms to rounding, also with a smaller
    active set (a dry start) and across a switch between the two forms."""
    cfg = synth.rain_basin(200, seed=5)
    N = cfg["dem"].shape[0] - 1
    nx = N + 4
    idx = [j * nx + i for i in range(60, 64) for j in range(40, 43)]
    kw = dict(n_const=cfg["n_const"], K=cfg["K"], psi=cfg["psi"], dtheta=cfg["dtheta"],
              flags=rift.TILE_ACCEL | rift.RAIN_AVG | rift.HYDROGRAPH | rift.MASS_BUDGET)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    th, vh = synth.series_hours_to_seconds(np.array([0.0, 0.02, 0.1], np.float32),
                                           np.array([0.0, 80.0, 20.0], np.float32))
    a = rift.grid_from_rasters(cfg["dem"], cfg["dx"], **kw)
    b = rift.grid_from_rasters(cfg["dem"], cfg["dx"], **kw)
    for g in (a, b):
        g.set_series(rift.SER_HYETO, t, v)
        g.set_source_cells(idx)
        g.set_series(rift.SER_HYDRO, th, vh)
        g.set_graphs(graphs)
    b.set_dt_kernel(True)
    for rep, n in enumerate((1, 16, 3, 16, 1, 19, 40)):
        if rep == 5:
            a.set_dt_kernel(True)    # switch forms mid-run: same trajectory
            b.set_dt_kernel(False)
        a.step(n)
        b.step(n)
        assert a.time == b.time, rep
        assert a.wet_cells() == b.wet_cells(), rep
        np.testing.assert_array_equal(a.rates(), b.rates())
        for f in ("depth", "hu", "hv", "F"):
            np.testing.assert_array_equal(a.field(f), b.field(f), err_msg=f"{f} after replay {rep}")
        wa, aa, na = a.tiles()
        wb, ab, nb = b.tiles()
        assert na == nb, rep
        np.testing.assert_array_equal(wa, wb)
        np.testing.assert_array_equal(aa, ab)
        # the budget terms are FP64 sums of the waves' atomics, in the order
        # the waves arrive: equal to rounding in the last bits
        np.testing.assert_allclose(a.budget(), b.budget(), rtol=1e-12, atol=1e-12)
    assert a.time[0] > 0.0 and a.wet_cells()[1] > 0
    a.close()
    b.close()


def test_launch_variants_bit_identical_and_autotuner(rift):
    """Every launch variant of the fused step (register budget per stage, work
    order) gives the default variant's state, dt and wet set bit for bit; the
    autotuner times them on a copy and leaves the handle's state untouched."""
    cfg = synth.rain_basin(160, seed=11)
    kw = dict(n_const=cfg["n_const"], K=cfg["K"], psi=cfg["psi"], dtheta=cfg["dtheta"],
              flags=rift.TILE_ACCEL | rift.RAIN_AVG)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    ref = None
    for var in range(rift.NVARIANTS):
        g = rift.grid_from_rasters(cfg["dem"], cfg["dx"], **kw)
        g.set_series(rift.SER_HYETO, t, v)
        g.set_variant(var)
        g.step(40)
        got = (g.time, [g.field(f) for f in ("depth", "hu", "hv", "F")], g.tiles()[0])
        if ref is None:
            ref = got
        else:
            assert got[0] == ref[0], var
            for a, b in zip(got[1], ref[1]):
                np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(got[2], ref[2])
        g.close()
    g = rift.grid_from_rasters(cfg["dem"], cfg["dx"], **kw)
    g.set_series(rift.SER_HYETO, t, v)
    g.step(20)
    before = (g.time, g.field("depth"))
    best, ms = g.autotune(16)
    # the fastest six of the sweep are timed twice more: the kept one is among them
    assert 0 <= best < rift.NVARIANTS and (ms > 0).all() and best in np.argsort(ms, kind="stable")[:6]
    assert g.time == before[0]
    np.testing.assert_array_equal(g.field("depth"), before[1])
    g.step(20)   # continues with the chosen variant
    g.close()


@pytest.mark.parametrize("accel", [True, False])
def test_device_volume_budget(rift, accel):
    """The device's FP64 volume budget (rift_budget, RIFT_MASS_BUDGET) closes
    to FP32 rounding of the state and matches the oracle's FP64 budget term
    by term (rain, source, drain, outflow through the outflow boundaries)."""
    n = 48
    N = n + 1
    dem = 6.0 * synth.fractal_dem(n, 0.6, 21) + np.linspace(0, 3, N)[None, :]
    h0 = np.maximum(3.5 - dem, 0.0)
    K = 5.0 * np.exp(0.5 * synth.normal(5, N * N).reshape(N, N))
    base_o = O.RAIN_AVG | O.DRAIN_AVG | O.HYDROGRAPH | (O.TILE_ACCEL if accel else 0)
    o = O.Oracle(dem, 2.0, h0=h0, n_const=0.03, K=K, psi=0.11, dtheta=0.3, flags=base_o)
    gflags = (rift.RAIN_AVG | rift.DRAIN_AVG | rift.HYDROGRAPH | rift.MASS_BUDGET |
              (rift.TILE_ACCEL if accel else 0))
    g = rift.grid_from_rasters(dem, 2.0, h0=h0, n_const=0.03, K=K, psi=0.11, dtheta=0.3,
                               flags=gflags)
    series = {O.SER_HYETO: ([0, 0.01, 0.05], [50, 120, 0]), O.SER_DRAIN: ([0, 0.05], [10, 10]),
              O.SER_HYDRO: ([0, 0.01], [20, 30])}
    idx = [j * o.nx + i for i in range(20, 23) for j in range(30, 32)]
    o.set_sources(idx)
    g.set_source_cells(idx)
    for which, (hours, vals) in series.items():
        hrs, vv = np.array(hours, np.float32), np.array(vals, np.float32)
        o.set_series_hours(which, hrs, vv)
        t, v = synth.series_hours_to_seconds(hrs, vv)
        g.set_series(which, t, v)
    for _ in range(4):
        o.step(40)
        g.step(40)
        bg, bo = g.budget(), o.budget()
        scale = np.abs(bg[:8]).max()
        assert abs(bg[8]) <= 1e-6 * scale, bg
    names = ["stored", "infiltrated", "v0", "rain", "source", "drain", "outflow"]
    tol = {"stored": 1e-4, "infiltrated": 1e-4, "v0": 1e-9, "rain": 1e-4, "source": 1e-4,
           "drain": 2e-3, "outflow": 2e-3}
    for k, name in enumerate(names):
        assert abs(bg[k] - bo[k]) <= tol[name] * max(abs(bo[k]), 1e-9), (name, bg[k], bo[k])
    # the exchange (dry clamps) is small here in both
    assert abs(bg[7] - bo[7]) <= 1e-3 * scale
    # rift_mass_balance (the reference's host sum of h and F, simulator.cu:
    # 609-637, as FP64 device sums): stored and infiltrated volume, and the
    # rain and source volume added, against the oracle
    mb = g.mass_balance()
    assert mb[0] == bg[0] and mb[1] == bg[1]
    for k, name in ((0, "stored"), (1, "infiltrated"), (2, "rain"), (3, "source")):
        ref = bo[{0: 0, 1: 1, 2: 3, 3: 4}[k]]
        assert abs(mb[k] - ref) <= 1e-4 * max(abs(ref), 1e-9), (name, mb[k], ref)
    g.close()
