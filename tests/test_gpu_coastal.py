"""The coastal-surge workload on the GPU: 8192 x 8192 (the benchmark's
headline, rift_amd.synth.coastal_surge(8192): surge maps every 600 s held in
device slots, hydrograph rectangle, Manning map, tile acceleration) and a
2048 x 2048 instance of the same generator.

8192^2, after the benchmark's spin-up to t = 3 h with the FP64 volume budget
on (the oracle cannot step this size in a test; these checks hold at any
size):
  - the device budget closes: |residual| <= 1e-4 of its largest term (FP32
    state over ~74,000 steps; the surge override and the outflow each move
    ~1.4e9 m^3);
  - depth is never negative;
  - every step, the active tile set equals the oracle's Track
    (grid.cu:1250-1276) and Grow (grid.cu:1188-1322) applied to the device's
    own wet flags, last active set and current surge map;
  - the surge window is the one InterpolatedGridSeries holds at that time
    and every map it needs is resident.

2048^2: the FP64 oracle continues from the GPU's state at t = 1800 s as a
hotstart (surge series restarted at that time with the maps around it; the
FP32 values given low bits, _generic) and
both advance side by side: dt within 5e-3 relative, depth within 1e-4 m
after the window, and tile flags equal except on thin-front tiles
(lockstep_lib).  Launch variants of the coastal instantiation stay
bit-identical.
"""
import numpy as np
import pytest

import bench
import oracle_lib as O
from lockstep_lib import check_tiles
from rift_amd import synth

pytestmark = pytest.mark.gpu


def _build(rift, n, budget=False):
    g, cfg = bench.build_coastal(n, budget=budget)
    return g, cfg


def test_coastal_8192_budget_positivity_and_tile_rule(rift):
    print("coastal 8192: build + spin-up to 3 h with the volume budget", flush=True)
    g, cfg = _build(rift, 8192, budget=True)
    steps = bench.spin_up(g, bench.SPINUP_S, label="test spin-up")
    b = g.budget()
    scale = np.abs(b[:8]).max()
    print(f"t = {g.time[0]:.1f} s after {steps} steps; budget residual {b[8]:.3e} of {scale:.3e}",
          flush=True)
    assert abs(b[8]) <= 1e-4 * scale, b
    assert b[6] > 1e8 and b[7] > 1e8            # outflow and surge exchange exercised
    g.set_budget(False)
    h = g.field("depth")
    assert h.min() >= 0.0
    wet_frac = (g.interior(h) > 1.19209e-07).mean()
    assert 0.05 <= wet_frac <= 0.25, wet_frac
    t, _ = g.time
    k0, k1, f, missing = g.map_state(rift.MAP_SURGE)
    assert (k0, k1) == (int(t // 600), int(t // 600) + 1) and missing == -1
    assert f == pytest.approx((t - 600 * k0) / 600.0, rel=1e-6)
    for k in range(4):
        wet, act, _ = g.tiles()
        surge = g.field("surge")                 # the window of the coming step
        _, want = O.track_and_grow(g.nbx, g.nby, wet, act, [(1, surge)])
        g.step(1)
        _, got_act, n_act = g.tiles()
        np.testing.assert_array_equal(got_act, want, err_msg=f"step {k}")
        assert n_act == int(want.sum())
    g.close()


def _lockstep_thin_front(o, g, nsteps):
    """Lockstep of the device and the FP64 oracle on a workload with a
    dry-land flood front: dt within 5e-3 relative, tile flags compared by
    lockstep_lib.check_tiles."""
    dt_err, nthin, prev = 0.0, 0, set()
    for k in range(nsteps):
        o.step()
        g.step()
        to, dto = o.time
        tg, dtg = g.time
        dt_err = max(dt_err, abs(dtg - dto) / dto)
        assert abs(dtg - dto) <= 5e-3 * dto, (k, dto, dtg)
        prev = check_tiles(g, o, prev, k)
        nthin += len(prev)
    return nsteps, dt_err, nthin


def _generic(a, rng):
    """FP32 values as FP64 with their low 32 bits filled: each nonzero value
    times (1 + u 2^-36), |u| <= 1.  The reference's sign-blind minmod
    (grid.cu:62-76) reads the low word of its double operands and returns 0
    where it is zero; FP32 values widened to double leave it zero, so an FP64
    hotstart from FP32 state would zero limiter slopes the reference's own
    FP64 state never zeroes.  The dither moves a 2 m depth by < 3e-11 m."""
    a = np.asarray(a, np.float64)
    return np.where(a != 0, a * (1 + rng.uniform(-1, 1, a.shape) * 2.0 ** -36), a)


def _spun_2048(rift, t_end=1800.0):
    g, cfg = _build(rift, 2048)
    bench.spin_up(g, t_end, chunk=1024)
    return g, cfg


def test_coastal_2048_lockstep_with_the_oracle(rift):
    print("coastal 2048: spin-up to 1800 s, then lockstep with the FP64 oracle", flush=True)
    g, cfg = _spun_2048(rift)
    t, dt = g.time
    o = O.Oracle(cfg["dem"], cfg["dx"], n=cfg["n"], flags=O.TILE_ACCEL | O.HYDROGRAPH | O.SURGE_GRID)
    o.set_sources(synth.source_cells(o.nx, cfg["source_rect"]))
    o.set_series_hours(O.SER_HYDRO, cfg["hydro_hours"], cfg["hydro_cfs"])
    rng = np.random.default_rng(2048)
    o.set_field("w", _generic(g.field("depth"), rng) + o.field("bc"))
    o.set_field("hu", _generic(g.field("hu"), rng))
    o.set_field("hv", _generic(g.field("hv"), rng))
    o.phase("apply_boundaries")
    wet, act, n_act = g.tiles()
    o.set_tiles(wet, np.where(act > 0, np.arange(g.ntiles), -np.arange(g.ntiles)))
    o.set_time(t, dt)
    k0 = g.map_state(rift.MAP_SURGE)[0]
    maps = np.stack([synth.surge_map(2048, k, cfg["surge_dt"]) for k in range(k0, k0 + 3)])
    o.set_map_series(O.MAP_SURGE, maps, cfg["surge_dt"], cfg["surge_tf"], first=k0)
    s_dev, s_orc = g.field("surge"), o.field("surge")
    inner = (slice(2, g.ny - 2), slice(2, g.nx - 2))
    np.testing.assert_allclose(s_dev[inner], s_orc[inner], rtol=0, atol=1e-6)
    assert 0.02 < n_act / g.ntiles < 0.5
    steps, dt_err, nthin = _lockstep_thin_front(o, g, 8)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    print(f"{steps} steps from t = {t:.1f} s: max|dh| {dh:.2e} m, dt rel {dt_err:.1e}, "
          f"thin-front tile flags differing: {nthin}", flush=True)
    assert dh <= 1e-4, dh
    assert g.interior(g.field("depth")).max() > 1.0     # the surge strip is flooded
    g.close()


def test_coastal_2048_launch_variants_bit_identical(rift):
    # g1 also forms dt in a kernel of its own (the surge window advances in
    # the dt tail): the same trajectory bit for bit
    g0, _ = _spun_2048(rift, 600.0)
    g1, _ = _spun_2048(rift, 600.0)
    assert g0.time == g1.time
    g0.set_variant(0)
    g1.set_variant(13)
    g1.set_dt_kernel(True)
    g0.step(24)
    g1.step(24)
    assert g0.time == g1.time
    for f in ("depth", "hu", "hv"):
        np.testing.assert_array_equal(g0.field(f), g1.field(f), err_msg=f)
    np.testing.assert_array_equal(g0.tiles()[1], g1.tiles()[1])
    g0.close()
    g1.close()
