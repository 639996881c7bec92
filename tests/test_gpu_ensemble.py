"""Batched ensembles (rift_ensemble_*): members of different sizes and
inputs advanced by one launch per phase give, member by member, exactly the
results of the same domain stepped alone (bit for bit: the same per-cell
arithmetic, only the dealing of work to waves differs), each with its own dt;
end times stop members independently.  Rain + Green-Ampt members run the
lean ensemble kernels; with summary maps (SAVE_MAX) the full ones; with the
FP64 volume budget the budget instantiations."""
import numpy as np
import pytest

from rift_amd import synth

pytestmark = pytest.mark.gpu


def _members():
    out = []
    for k, (n, m, dx, peak, nman) in enumerate([(64, 80, 10.0, 60.0, 0.03), (100, 60, 5.0, 90.0, 0.05),
                                                  (40, 40, 10.0, 40.0, 0.02), (128, 20, 8.0, 75.0, 0.04)]):
        size = max(n, m)
        dem = 40.0 * synth.fractal_dem(size, 0.6, 30 + k)[: m + 1, : n + 1]
        N = (m + 1) * (n + 1)
        K = 5.0 * np.exp(0.8 * synth.normal(40 + k, N).reshape(m + 1, n + 1))
        hours = np.array([0.0, 0.02, 0.05], np.float32)
        mmhr = np.array([0.0, peak, 0.0], np.float32)
        out.append({"dem": dem, "dx": dx, "n_const": nman, "K": K, "psi": 0.11, "dtheta": 0.3,
                    "hours": hours, "mmhr": mmhr})
    return out


def _wet_members():
    # heavier, longer rain on less permeable ground: every member holds
    # 0.04-0.25 m of water by step 150 and 0.15-0.27 m by step 500
    out = []
    for k, (n, m, dx, peak, nman) in enumerate([(64, 80, 10.0, 150.0, 0.03), (100, 60, 5.0, 200.0, 0.05),
                                                  (40, 40, 10.0, 120.0, 0.02), (128, 20, 8.0, 180.0, 0.04)]):
        size = max(n, m)
        dem = 40.0 * synth.fractal_dem(size, 0.6, 30 + k)[: m + 1, : n + 1]
        N = (m + 1) * (n + 1)
        K = 1.0 * np.exp(0.8 * synth.normal(40 + k, N).reshape(m + 1, n + 1))
        hours = np.array([0.0, 0.05, 0.25], np.float32)
        mmhr = np.array([0.0, peak, 0.0], np.float32)
        out.append({"dem": dem, "dx": dx, "n_const": nman, "K": K, "psi": 0.11, "dtheta": 0.3,
                    "hours": hours, "mmhr": mmhr})
    return out


def _flags(rift, extra=0):
    return rift.TILE_ACCEL | rift.RAIN_AVG | extra


def _single(rift, m, extra=0):
    g = rift.grid_from_rasters(m["dem"], m["dx"], n_const=m["n_const"], K=m["K"], psi=m["psi"],
                               dtheta=m["dtheta"], flags=_flags(rift, extra))
    t, v = synth.series_hours_to_seconds(m["hours"], m["mmhr"])
    g.set_series(rift.SER_HYETO, t, v)
    return g


def _ensemble(rift, members, extra=0):
    spec = [dict(dem=m["dem"], dx=m["dx"], n_const=m["n_const"], K=m["K"], psi=m["psi"],
                 dtheta=m["dtheta"], flags=_flags(rift, extra)) for m in members]
    e = rift.Ensemble(spec)
    for g, m in zip(e.members, members):
        t, v = synth.series_hours_to_seconds(m["hours"], m["mmhr"])
        g.set_series(rift.SER_HYETO, t, v)
    return e


@pytest.mark.parametrize("extra", ["none", "SAVE_MAX", "MASS_BUDGET"])
def test_ensemble_members_match_single_domain_runs(rift, extra):
    members = _members()
    print("ensemble of", len(members), extra, flush=True)
    xf = 0 if extra == "none" else getattr(rift, extra)
    e = _ensemble(rift, members, xf)
    singles = [_single(rift, m, xf) for m in members]
    for chunk in (1, 7, 40, 32):
        e.step(chunk)
        for g in singles:
            g.step(chunk)
        times = set()
        for i, (ge, gs) in enumerate(zip(e.members, singles)):
            assert ge.time == gs.time, (i, ge.time, gs.time)
            times.add(ge.time[1])
            for f in ("depth", "hu", "hv", "F") + (("h_max", "q_max") if extra == "SAVE_MAX" else ()):
                np.testing.assert_array_equal(ge.field(f), gs.field(f), err_msg=f"{i} {f}")
            np.testing.assert_array_equal(ge.tiles()[0], gs.tiles()[0])
            np.testing.assert_array_equal(ge.tiles()[1], gs.tiles()[1])
    assert len(times) > 1   # each member has its own dt
    for g in singles:
        g.close()
    e.close()


def test_ensemble_end_times(rift):
    members = _members()[:3]
    e = _ensemble(rift, members)
    tfs = [30.0, 75.0, 50.0]
    for i, tf in enumerate(tfs):
        e.set_tf(i, tf)
    steps = e.run(check_every=8)
    assert steps > 0 and e.finished() == 3
    ends = []
    for g, tf in zip(e.members, tfs):
        t, dt = g.time
        assert t >= tf and t - dt < tf, (t, dt, tf)
        ends.append((t, g.field("depth")))
    e.step(5)   # ended members stay as they are
    for g, (t, h) in zip(e.members, ends):
        assert g.time[0] == t
        np.testing.assert_array_equal(g.field("depth"), h)
    e.close()


def test_ended_members_leave_the_launches(rift):
    """Members seen at their end time (rift_ensemble_finished) are no longer
    launched: the members still running match their single-domain runs bit
    for bit while the others have left, and each ended member holds the
    state of its single run at its last step."""
    members = _members()
    e = _ensemble(rift, members)
    singles = [_single(rift, m) for m in members]
    tfs = [20.0, 1e9, 35.0, 1e9]
    for i, tf in enumerate(tfs):
        e.set_tf(i, tf)
    steps = 0
    while e.finished() < 2:
        e.step(4)
        steps += 4
        for g, tf in zip(singles, tfs):   # a single run stops where its member does
            for _ in range(4):
                if g.time[0] < tf:
                    g.step(1)
    assert steps < 400
    for _ in range(3):   # two members stepped, two left out
        e.step(9)
        for k in (1, 3):
            singles[k].step(9)
        assert e.finished() == 2
    for i, (ge, gs) in enumerate(zip(e.members, singles)):
        assert ge.time == gs.time, (i, ge.time, gs.time)
        if tfs[i] < 1e9:
            t, dt = ge.time
            assert t >= tfs[i] and t - dt < tfs[i], (i, t, dt)
        for f in ("depth", "hu", "hv", "F"):
            np.testing.assert_array_equal(ge.field(f), gs.field(f), err_msg=f"{i} {f}")
        np.testing.assert_array_equal(ge.tiles()[0], gs.tiles()[0])
    for g in singles:
        g.close()
    e.close()


def test_ensemble_of_more_than_64_members(rift):
    """Members are located 64 at a time (one per lane); 70 small members of
    varied sizes, some dense and some with few active tiles, still match
    their single-domain runs bit for bit."""
    members = []
    for k in range(70):
        n, m = 20 + (k * 7) % 37, 18 + (k * 11) % 29
        dem = 15.0 * synth.fractal_dem(max(n, m), 0.6, 300 + k)[: m + 1, : n + 1]
        K = 4.0 * np.exp(0.5 * synth.normal(400 + k, (m + 1) * (n + 1)).reshape(m + 1, n + 1))
        members.append({"dem": dem, "dx": 5.0, "n_const": 0.03 + 0.0005 * k, "K": K, "psi": 0.11,
                        "dtheta": 0.3, "hours": np.array([0.0, 0.01, 0.03], np.float32),
                        "mmhr": np.array([0.0, 40.0 + k, 0.0], np.float32)})
    print("ensemble of", len(members), flush=True)
    e = _ensemble(rift, members)
    singles = [_single(rift, m) for m in members]
    for chunk in (3, 29):
        e.step(chunk)
        for g in singles:
            g.step(chunk)
        for i, (ge, gs) in enumerate(zip(e.members, singles)):
            assert ge.time == gs.time, (i, ge.time, gs.time)
            for f in ("depth", "hu", "hv", "F"):
                np.testing.assert_array_equal(ge.field(f), gs.field(f), err_msg=f"{i} {f}")
    for g in singles:
        g.close()
    e.close()



def test_ensemble_members_match_the_fp64_oracle(rift):
    """Every member of an ensemble advanced together against the FP64 oracle
    of the same domain stepped alone: dt per step within 5e-3 relative (thin
    sheet flow over rough terrain, as in the single-domain rain-basin parity
    test), tile flags compared as where a flood front crosses dry land
    (lockstep_lib: they may differ only on tiles whose wettest cell is
    below 1e-5 m in both), depth and infiltrated depth within 1e-4 m after
    500 steps (t = 350-560 s, every member wet, 0.15-0.27 m deep)."""
    import oracle_lib as O
    from lockstep_lib import check_tiles
    members = _wet_members()
    e = _ensemble(rift, members)
    oracles = []
    for m in members:
        o = O.Oracle(m["dem"], m["dx"], n_const=m["n_const"], K=m["K"], psi=m["psi"],
                     dtheta=m["dtheta"], flags=O.TILE_ACCEL | O.RAIN_AVG)
        o.set_series_hours(O.SER_HYETO, m["hours"], m["mmhr"])
        oracles.append(o)
    dts = set()
    prev = [set() for _ in members]
    nthin = 0
    for k in range(500):
        e.step(1)
        for i, (g, o) in enumerate(zip(e.members, oracles)):
            o.step()
            (tg, dtg), (to, dto) = g.time, o.time
            assert abs(dtg - dto) <= 5e-3 * dto, (k, i, dtg, dto)
            dts.add(dtg)
            prev[i] = check_tiles(g, o, prev[i], (k, i))
            nthin += len(prev[i])
    print(f"thin-front tile flags differing over the run: {nthin}", flush=True)
    assert len(dts) > len(members)
    for i, (g, o) in enumerate(zip(e.members, oracles)):
        dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
        dF = np.abs(g.interior(g.field("F")) - o.interior(o.field("F"))).max()
        hmax = g.interior(g.field("depth")).max()
        print(f"member {i}: t = {g.time[0]:.1f} s, max depth {hmax:.3f} m, max|dh| {dh:.2e} m, "
              f"max|dF| {dF:.2e} m", flush=True)
        assert hmax > 0.1, (i, hmax)
        assert dh <= 1e-4 and dF <= 1e-4, (i, dh, dF)
    e.close()


def test_full_size_ensemble_properties(rift):
    """The benchmark's ensemble at full size (synth.ensemble_scenarios(64, 4):
    sides 512-2048, 102.8 M cells, rain + Green-Ampt + Manning, tf = 1 h)
    run to its end with FP64 volume budgets: every member reaches its end
    time and closes its budget to 1e-4 of its largest term with depth >= 0;
    the two smallest members equal their single-domain runs (rift_step to
    the same end time) bit for bit."""
    scen = synth.ensemble_scenarios(64, 4)
    flags = rift.TILE_ACCEL | rift.RAIN_AVG | rift.MASS_BUDGET
    spec = [dict(dem=m["dem"], dx=m["dx"], n_const=m["n_const"], K=m["K"], psi=m["psi"],
                 dtheta=m["dtheta"], flags=flags) for m in scen]
    e = rift.Ensemble(spec)
    for i, (g, m) in enumerate(zip(e.members, scen)):
        t, v = synth.series_hours_to_seconds(m["hours"], m["mmhr"])
        g.set_series(rift.SER_HYETO, t, v)
        e.set_tf(i, m["tf"])
    print(f"ensemble of {len(scen)}: {sum(m['side'] ** 2 for m in scen) / 1e6:.1f} M cells", flush=True)
    steps = e.run(check_every=64)
    print(f"ran {steps} ensemble steps", flush=True)
    assert e.finished() == len(scen)
    worst = 0.0
    for i, (g, m) in enumerate(zip(e.members, scen)):
        assert g.time[0] >= m["tf"], (i, g.time, m["tf"])
        b = g.budget()
        rel = abs(b[8]) / np.abs(b[:8]).max()
        worst = max(worst, rel)
        assert rel <= 1e-4, (i, b)
        assert g.field("depth").min() >= 0.0, i
    print(f"largest budget residual {worst:.2e} of its largest term", flush=True)
    smallest = sorted(range(len(scen)), key=lambda i: scen[i]["side"])[:2]
    for i in smallest:
        m = scen[i]
        g = rift.grid_from_rasters(m["dem"], m["dx"], n_const=m["n_const"], K=m["K"], psi=m["psi"],
                                   dtheta=m["dtheta"], flags=flags)
        t, v = synth.series_hours_to_seconds(m["hours"], m["mmhr"])
        g.set_series(rift.SER_HYETO, t, v)
        n = 0
        while g.time[0] < m["tf"]:
            g.step(1)
            n += 1
        ge = e.members[i]
        print(f"member {i} (side {m['side']}): {n} steps alone", flush=True)
        assert ge.time == g.time, (i, ge.time, g.time)
        for f in ("depth", "hu", "hv", "F"):
            np.testing.assert_array_equal(ge.field(f), g.field(f), err_msg=f"member {i} {f}")
        np.testing.assert_array_equal(ge.tiles()[1], g.tiles()[1])
        g.close()
    e.close()
