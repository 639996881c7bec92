"""The device-resident surge series (rift_map_series_*) on the GPU.

The reference keeps an interpolated surge series as two device planes, an
InterpGrid launch per step and a host p_update after every step
(GridSeries.cu:291-384, simulator.cu:827-829).  The device series holds the
maps in slots, advances p_update's window in the dt kernel and interpolates
per cell where the step reads surge.  The same series driven from the host
through map planes 1 and 2 and rift_interp_map (the plane-0 path, kept for
callers that load maps themselves) gives the same FP32 values, so the two
runs must be bit-identical: t, dt, tile sets every step and the fields at
the end.  The per-phase path materialises the window into plane 0 and is
checked against the FP64 oracle (dt 5e-3 relative, depth 1e-4 m, tile sets
identical every step).
"""
import numpy as np
import pytest

import oracle_lib as O
from rift_amd import synth

pytestmark = pytest.mark.gpu

ND = -9999.0


def _case(n=64, nmaps=6, surge_dt=15, surge_tf=75.0):
    """A small coastal case whose surge maps exercise every branch of the
    nodata rule and both signs: the strip widens on odd maps (cells with data
    in one plane only), and map 2 holds a negative surge on its west half."""
    c = synth.coastal_surge(n, nmaps=nmaps)
    N = n + 1
    maps = np.full((nmaps, N, N), ND)
    for k in range(nmaps):
        rows = 6 if k % 2 else 4
        maps[k, :rows, :] = 1.5 + 0.6 * k
    maps[2, :4, : N // 2] = -0.5
    c["surge_maps"], c["surge_dt"], c["surge_tf"] = maps, surge_dt, surge_tf
    c["hydro_hours"] = np.array([0.0, 0.005, 0.02], np.float32)
    c["hydro_cfs"] = np.array([0.0, 400.0, 300.0], np.float32)
    return c


def _grid(rift, c, per_phase=False):
    flags = rift.TILE_ACCEL | rift.SURGE_GRID | rift.HYDROGRAPH
    g = rift.grid_from_rasters(c["dem"], c["dx"], n=c["n"], flags=flags, per_phase=per_phase)
    t, v = synth.series_hours_to_seconds(c["hydro_hours"], c["hydro_cfs"])
    g.set_series(rift.SER_HYDRO, t, v)
    g.set_source_cells(synth.source_cells(g.nx, c["source_rect"]))
    return g


def _cells(rift, c):
    return [rift.map_cells(m, 1.0, True, ND) for m in c["surge_maps"]]


class HostInterp:
    """InterpolatedGridSeries::p_update (GridSeries.cu:325-384) driven from the
    host: maps into planes 1 and 2, plane 0 by rift_interp_map each step."""

    def __init__(self, rift, g, cells, dt, tmax):
        self.rift, self.g, self.cells, self.dt, self.tmax = rift, g, cells, dt, tmax
        fill = np.full((g.ny, g.nx), ND)
        for p in range(3):
            g.set_map(rift.MAP_SURGE, p, fill)
        self.in_time, self.next, self.done = -9999.0, 0.0, False

    def update(self, t):
        g, S = self.g, self.rift.MAP_SURGE
        if self.in_time < 0:
            idx = int(np.trunc(t / self.dt))
            g.set_map(S, 1, self.cells[idx])
            self.in_time = float(idx * self.dt)
            self.next = self.in_time + self.dt
            g.set_map(S, 2, self.cells[idx + 1])
        elif t >= self.tmax:
            if not self.done:
                self.in_time = self.next
                self.next = self.in_time + self.dt
                g.copy_map(S, 1, 2)
            self.done = True
        elif t >= self.next:
            g.copy_map(S, 1, 2)
            self.in_time = self.next
            self.next = self.in_time + self.dt
            g.set_map(S, 2, self.cells[int(np.trunc(self.next / self.dt))])
        g.interp_map(S, (t - self.in_time) / (self.next - self.in_time))


class LazyLoader:
    """The host side of a two-slot device series: the same window, used only
    to load each map when p_update would read it (rift_sim's surge series)."""

    def __init__(self, rift, g, cells, dt, tmax, t0=0.0):
        self.rift, self.g, self.cells, self.dt, self.tmax = rift, g, cells, dt, tmax
        g.map_series(rift.MAP_SURGE, 2, dt, tmax)
        idx = int(np.trunc(t0 / dt))
        g.load_map(rift.MAP_SURGE, idx, cells[idx])
        g.load_map(rift.MAP_SURGE, idx + 1, cells[idx + 1])
        self.in_time, self.next, self.done = float(idx * dt), float(idx * dt + dt), False
        g.start_maps(rift.MAP_SURGE, t0)

    def update(self, t):
        if t >= self.tmax:
            if not self.done:
                self.in_time = self.next
                self.next = self.in_time + self.dt
            self.done = True
        elif t >= self.next:
            self.in_time = self.next
            self.next = self.in_time + self.dt
            k = int(np.trunc(self.next / self.dt))
            self.g.load_map(self.rift.MAP_SURGE, k, self.cells[k])


def test_device_series_bit_identical_to_host_interpolation(rift):
    c = _case()
    cells = _cells(rift, c)
    host = _grid(rift, c)
    hi = HostInterp(rift, host, cells, c["surge_dt"], c["surge_tf"])
    hi.update(0.0)
    resident = _grid(rift, c)
    resident.map_series(rift.MAP_SURGE, len(cells), c["surge_dt"], c["surge_tf"])
    for k, m in enumerate(cells):
        resident.load_map(rift.MAP_SURGE, k, m)
    resident.start_maps(rift.MAP_SURGE, 0.0)
    lazy = _grid(rift, c)
    ll = LazyLoader(rift, lazy, cells, c["surge_dt"], c["surge_tf"])
    print("surge series: host / resident / two-slot runs to t = 90 s", flush=True)
    steps, t = 0, 0.0
    while t < 90.0:
        host.step(1)
        t, dt = host.time
        hi.update(t)
        resident.step(1)
        lazy.step(1)
        ll.update(lazy.time[0])
        steps += 1
        for g in (resident, lazy):
            assert g.time == (t, dt), (steps, g.time, t, dt)
            gw, ga, _ = g.tiles()
            hw, ha, _ = host.tiles()
            np.testing.assert_array_equal(gw, hw)
            np.testing.assert_array_equal(ga, ha)
    assert steps > 100
    for g in (resident, lazy):
        for f in ("depth", "hu", "hv", "surge"):
            np.testing.assert_array_equal(g.field(f), host.field(f), err_msg=f)
        k0, k1, _, missing = g.map_state(rift.MAP_SURGE)
        assert missing == -1 and k0 == k1 == 5   # frozen on the last map past surge_tf
    # the strip is wet and the surge reached inland
    h = host.interior(host.field("depth"))
    assert (h[:3] > 1.0).all()
    for g in (host, resident, lazy):
        g.close()


def test_device_series_reports_a_map_not_loaded(rift):
    """With two slots and no host loads, the window that needs map 2 finds
    slot 0 still holding map 0: the step reads in-bounds stale data and the
    first missing index is reported."""
    c = _case()
    cells = _cells(rift, c)
    g = _grid(rift, c)
    g.map_series(rift.MAP_SURGE, 2, c["surge_dt"], c["surge_tf"])
    g.load_map(rift.MAP_SURGE, 0, cells[0])
    g.load_map(rift.MAP_SURGE, 1, cells[1])
    g.start_maps(rift.MAP_SURGE, 0.0)
    while g.time[0] < 14.0:
        g.step(1)
    assert g.map_state(rift.MAP_SURGE)[3] == -1
    while g.time[0] < 16.0:
        g.step(1)
    k0, k1, f, missing = g.map_state(rift.MAP_SURGE)
    assert (k0, k1) == (1, 2) and 0.0 <= f < 1.0
    assert missing == 2
    g.close()


@pytest.mark.parametrize("per_phase", [False, True], ids=["fused", "per_phase"])
def test_device_series_against_oracle(rift, per_phase):
    c = _case()
    cells = _cells(rift, c)
    g = _grid(rift, c, per_phase=per_phase)
    ll = LazyLoader(rift, g, cells, c["surge_dt"], c["surge_tf"])
    o = O.Oracle(c["dem"], c["dx"], n=c["n"], flags=O.TILE_ACCEL | O.HYDROGRAPH | O.SURGE_GRID)
    o.set_sources(synth.source_cells(o.nx, c["source_rect"]))
    o.set_series_hours(O.SER_HYDRO, c["hydro_hours"], c["hydro_cfs"])
    o.set_map_series(O.MAP_SURGE, c["surge_maps"], c["surge_dt"], c["surge_tf"])
    print(f"surge series vs oracle, per_phase={per_phase}", flush=True)
    steps = 0
    while o.time[0] < 90.0:
        o.step(1)
        g.step(1)
        tg, dtg = g.time
        ll.update(tg)
        to, dto = o.time
        assert abs(dtg - dto) <= 5e-3 * dto, (steps, dtg, dto)
        gw, ga, _ = g.tiles()
        ow, oa, _ = o.tiles()
        np.testing.assert_array_equal(gw, ow)
        assert set(np.nonzero(ga)[0].tolist()) == set(int(v) for v in oa if v > -1) | {0}, steps
        steps += 1
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    s = g.field("surge")
    so = o.field("surge")
    inner = (slice(2, g.ny - 2), slice(2, g.nx - 2))
    np.testing.assert_allclose(s[inner], so[inner], rtol=0, atol=1e-6)
    g.close()
