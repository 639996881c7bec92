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

