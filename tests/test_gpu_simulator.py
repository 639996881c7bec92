"""The Simulator driver end to end on the GPU against the oracle.

A generated workload is written as the files a RIFT run reads (keyword file,
ESRI-ASCII rasters, CSV series, `-###` map series); the product opens it
through the C ABI (rift_sim_open: ReadUserParams + InitSimulation,
simulator.cu:25-493), steps it with RunSimulation's sequence (device step,
then the gridded-series updates and periodic output, simulator.cu:774-850)
and writes the reference's output files.  The oracle is given the same
arrays directly and is stepped in lockstep.

Tolerances as in test_gpu_parity: dt relative 5e-3 (thin sheet flow on rough
terrain), depth 1e-4 m.  Hotstart restarts are compared with the
uninterrupted run of the product itself.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from rift_amd import synth

pytestmark = pytest.mark.gpu

DT_RTOL = 5e-3
DH_TOL = 1e-4


def _coastal_case(n=48):
    c = synth.coastal_surge(n, nmaps=6)
    N = n + 1
    strip = 4
    maps = np.full((6, N, N), -9999.0)
    for k in range(6):
        maps[k, :strip, :] = 1.5 + 0.6 * k          # surge depth rising every map
    c["surge_maps"], c["surge_dt"], c["surge_tf"] = maps, 15, 75.0
    c["K"] = 5.0 * np.exp(0.8 * synth.normal(41, N * N).reshape(N, N))
    c["psi"], c["dtheta"] = 0.11, 0.3
    c["hyeto_hours"] = np.array([0.0, 0.01, 0.02], np.float32)
    c["hyeto_mmhr"] = np.array([0.0, 60.0, 20.0], np.float32)
    c["hydro_hours"] = np.array([0.0, 0.005, 0.02], np.float32)
    c["hydro_cfs"] = np.array([0.0, 400.0, 300.0], np.float32)
    c["source_rect"] = (20, 23, 30, 32)
    c["tf"] = 45.0
    return c


def _oracle_for(c, flags):
    o = O.Oracle(c["dem"], c["dx"], n=c.get("n"), n_const=c.get("n_const", 0.0), K=c.get("K"),
                 psi=c.get("psi", 0.0), dtheta=c.get("dtheta", 0.0), flags=flags)
    if "hyeto_hours" in c:
        o.set_series_hours(O.SER_HYETO, c["hyeto_hours"], c["hyeto_mmhr"])
    if "hydro_hours" in c:
        i0, i1, j0, j1 = c["source_rect"]
        o.set_sources([j * o.nx + i for i in range(i0, i1 + 1) for j in range(j0, j1 + 1)])
        o.set_series_hours(O.SER_HYDRO, c["hydro_hours"], c["hydro_cfs"])
    if "surge_maps" in c:
        o.set_map_series(O.MAP_SURGE, c["surge_maps"], c["surge_dt"], c["surge_tf"])
    if "rain_maps" in c:
        o.set_map_series(O.MAP_RAIN, c["rain_maps"], c["rain_dt"], c["rain_tf"])
    if "drain_maps" in c:
        o.set_map_series(O.MAP_DRAIN, c["drain_maps"], c["drain_dt"], c["drain_tf"])
    return o


def _lockstep(rift, cfg_path, o, dt_print=None):
    """Steps the product's Simulation and the oracle together; returns the
    simulation, the oracle depths at the product's print steps and the step
    count."""
    sim = rift.Simulation(cfg_path)
    t, _, tf = sim.time()
    prints, t_print, steps = [], 0.0, 0
    while t < tf:
        t = sim.step()
        o.step(1)
        steps += 1
        ot, _ = o.time
        assert abs(t - ot) <= DT_RTOL * ot, (steps, t, ot)
        # wet and active tile sets (Grow with TrackRainfall/TrackSurge) are
        # integer work: identical every step
        gw, ga, _ = sim.grid.tiles()
        ow, oa, _ = o.tiles()
        np.testing.assert_array_equal(gw, ow, err_msg=f"wet tiles, step {steps}")
        assert set(np.nonzero(ga)[0].tolist()) == set(int(v) for v in oa if v > -1), steps
        if dt_print is not None and t > t_print:
            prints.append(o.interior(o.depth()).copy())
            t_print += dt_print
    return sim, prints, steps


def test_simulator_coastal_surge_end_to_end(rift, tmp_path, monkeypatch):
    c = _coastal_case()
    case_dir = tmp_path / "case"
    cfg = synth.write_case(c, str(case_dir), save_max=1, save_arrival=1, h_print=1, q_print=1,
                           dt_print=15)
    monkeypatch.chdir(case_dir)
    flags = (O.TILE_ACCEL | O.HYDROGRAPH | O.RAIN_AVG | O.SURGE_GRID | O.SAVE_MAX |
             O.SAVE_ARRIVAL)
    o = _oracle_for(c, flags)
    print("simulator coastal case", cfg, flush=True)
    sim, prints, steps = _lockstep(rift, cfg, o, dt_print=15.0)
    assert steps > 50 and len(prints) >= 3
    g = sim.grid
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= DH_TOL, dh
    # the surge strip is wet and water reached past it
    assert (g.interior(g.field("depth"))[:3] > 1.0).all()
    for k, want in enumerate(prints):
        got, hdr = rift.read_raster(str(case_dir / f"h{k}.txt"))
        assert hdr["cellsize"] == c["dx"] and hdr["xllcorner"] == 0.5 * c["dx"]
        assert np.abs(got - want).max() <= DH_TOL, k
        assert (case_dir / f"q{k}.txt").exists()
    sim.close()
    hmax, _ = rift.read_raster(str(case_dir / "peak_flood_depth.txt"))
    assert np.abs(hmax - o.interior(o.field("h_max"))).max() <= DH_TOL
    twet, _ = rift.read_raster(str(case_dir / "flood_wave_arrival.txt"))
    ow = o.interior(o.field("t_wet"))
    close = np.abs(twet - ow) <= 0.5
    assert close.mean() >= 0.99, close.mean()
    for name in ("peak_unit_flow.txt", "time_to_peak.txt", "time_to_dry.txt"):
        assert (case_dir / name).exists()

    # the command-line driver writes the same summary maps (same code, same
    # device: bit-identical)
    cli_cfg = case_dir / "cli.cfg"
    text = open(cfg).read().replace("output = .", "output = cli_out")
    cli_cfg.write_text(text.replace("h_print = 1", "h_print = 0").replace("q_print = 1", "q_print = 0"))
    (case_dir / "cli_out").mkdir()
    exe = os.path.join(os.path.dirname(rift.LIB_PATH), "rift")
    print("simulator CLI run", flush=True)
    r = subprocess.run([exe, str(cli_cfg)], cwd=case_dir, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    a = (case_dir / "peak_flood_depth.txt").read_bytes()
    b = (case_dir / "cli_out" / "peak_flood_depth.txt").read_bytes()
    assert a == b


@pytest.mark.parametrize("interp", [False, True])
def test_simulator_gridded_rain_and_drain(rift, tmp_path, monkeypatch, interp):
    """Held (HyetographGridSeries) and interpolated rain maps plus held drain
    maps, with the four-corner nodata rule of p_interp (nnd > 2)."""
    n = 40
    N = n + 1
    c = {"dem": 8.0 * synth.fractal_dem(n, 0.6, 13), "dx": 4.0, "n_const": 0.03, "tf": 200.0}
    nm = 7
    rain = np.empty((nm, N, N))
    for k in range(nm):
        rain[k] = 40.0 + 60.0 * synth.uniform(100 + k, N * N).reshape(N, N) + 15.0 * k
    rain[:, 10:14, 5:9] = 0.0
    drain = np.full((nm, N, N), 5.0)
    drain[:, 20:, :] = 25.0
    c.update(rain_maps=rain, rain_dt=50, rain_tf=230.0, drain_maps=drain, drain_dt=50,
             drain_tf=230.0)
    case_dir = tmp_path / "rain"
    case_dir.mkdir()
    for k in range(nm):
        synth._write_asc(str(case_dir / ("rain-%03d.txt" % k)), rain[k], c["dx"])
        synth._write_asc(str(case_dir / ("drain-%03d.txt" % k)), drain[k], c["dx"])
    kw = dict(hyetograph_prefix="rain", hyetograph_dt=50, hyetograph_tf=230.0,
              drain_prefix="drain", drain_dt=50, drain_tf=230.0)
    if interp:
        kw["hyetograph_interp"] = 1
    cfg = synth.write_case(c, str(case_dir), **kw)
    monkeypatch.chdir(case_dir)
    flags = O.TILE_ACCEL | O.RAIN_GRID | O.DRAIN_GRID | (O.RAIN_INTERP if interp else 0)
    o = _oracle_for(c, flags)
    print("simulator gridded rain interp=%d" % interp, flush=True)
    sim, _, steps = _lockstep(rift, cfg, o)
    g = sim.grid
    assert steps >= 10
    # held maps are equal up to float storage; interpolated ones also move
    # with the small difference of the two clocks (the time factor)
    t_gpu, _, _ = sim.time()
    t_orc, _ = o.time
    slope = np.abs(np.diff(rain, axis=0)).max() / 3.6e6 / 50.0
    for f in ("rainmap", "drainmap"):
        a = g.interior(g.field(f))
        b = o.interior(o.field(f))
        tol = 1e-6 * max(np.abs(b).max(), 1e-12) + 1e-15
        if interp and f == "rainmap":
            tol += slope * abs(t_gpu - t_orc)
        assert np.abs(a - b).max() <= tol, f
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= DH_TOL, dh
    sim.close()


def _restart_case():
    n = 48
    N = n + 1
    c = synth.rain_basin(n, seed=5)
    c["hyeto_hours"] = np.array([0.0, 0.01, 0.03], np.float32)
    c["hyeto_mmhr"] = np.array([30.0, 90.0, 10.0], np.float32)
    c["tf"] = 1.0e9
    assert c["dem"].shape == (N, N)
    return c


def test_hotstart_restart_equivalence(rift, tmp_path, monkeypatch):
    """WriteHotstart / ReadHotstart (simulator.cu:912-1090): a run restarted
    from a hotstart continues as the uninterrupted run (rain on every tile
    keeps every tile active, so no stale tile maxima are lost)."""
    c = _restart_case()
    case_dir = tmp_path / "hot"
    cfg = synth.write_case(c, str(case_dir))
    monkeypatch.chdir(case_dir)
    print("hotstart: first run", flush=True)
    a = rift.Simulation(cfg)
    for _ in range(30):
        a.step()
    a.write_hotstart(str(case_dir / "hot.bin"))
    t_hot, dt_hot, _ = a.time()
    for _ in range(30):
        a.step()
    ta, _, _ = a.time()
    ha = a.grid.field("depth")
    Fa = a.grid.field("F")
    a.close()

    nx, ny = c["dem"].shape[1] + 3, c["dem"].shape[0] + 3
    nbx, nby = -(-nx // 16), -(-ny // 14)
    gs = (-(-nbx // 16) * 16) * (-(-nby // 14) * 14)
    raw = (case_dir / "hot.bin").read_bytes()
    assert len(raw) == 16 + 82 + 4 * nx * ny * 8 + gs + 4 * gs
    t0, dt0 = np.frombuffer(raw[:16], np.float64)
    assert t0 == t_hot and dt0 == dt_hot
    assert np.frombuffer(raw[16:24], np.int32).tolist() == [nx - 3, ny - 3]

    print("hotstart: restarted run", flush=True)
    with open(cfg, "a") as f:
        f.write("read_hotstart = hot.bin\n")
    b = rift.Simulation(cfg)
    tb0, dtb0, _ = b.time()
    assert tb0 == t_hot and dtb0 == dt_hot
    for _ in range(30):
        b.step()
    tb, _, _ = b.time()
    assert abs(tb - ta) <= 1e-9 * ta, (ta, tb)
    assert np.abs(b.grid.field("depth") - ha).max() <= 1e-6
    assert np.abs(b.grid.field("F") - Fa).max() <= 1e-7
    b.close()


def test_simulator_reports_bad_inputs(rift, tmp_path, monkeypatch):
    """Missing files and wrong raster sizes are errors with a message (the
    reference throws or exits)."""
    c = _coastal_case(16)
    case_dir = tmp_path / "bad"
    cfg = synth.write_case(c, str(case_dir))
    monkeypatch.chdir(case_dir)
    os.remove(case_dir / "hydrograph.csv")
    with pytest.raises(rift.RiftError, match="hydrograph.csv"):
        rift.Simulation(cfg)
    cfg2 = synth.write_case(c, str(case_dir), name="bad2")
    synth._write_asc(str(case_dir / "n.asc"), np.full((5, 5), 0.03), c["dx"])
    with pytest.raises(rift.RiftError, match="incorrect size"):
        rift.Simulation(cfg2)
