"""Pins of the oracle's map series, tile bookkeeping, geometry, summary maps
and dt against what the reference's launch geometry or the mathematics
fixes (the reference ships no vectors for any of them).  Each check is chosen
so that a plausible slip fails:
  TrackRainfall_k/TrackSurge_k  the 16-row window stride (grid.cu:1239, 1265)
                                vs the 14-row tile rows, by brute force over
                                every single-cell map of a tile grid
  p_interp                      all 16 nodata patterns of a cell's corners,
                                `nnd > 2` (GridSeries.cu:91-122)
  map timing                    interpolated surge linear in time, frozen at
                                surge_tf; held rain from the map at t + dt
                                (GridSeries.cu:144-178, 261-283, 325-384)
  surge override                h* = surge - dU dt on a raining flat tank
                                (grid.cu:894-904)
  tile wet flags                thread (0,0)'s rule: a tile whose origin cell
                                is a ghost keeps its flag (grid.cu:1160-1170)
  geometry                      linear DEM: face beds and BC are the plane's
                                values at face and cell centres, ghost faces
                                repeat interior ones (simulator.cu:384-431)
  summary maps                  h_max (float-rounded by fmaxf), t_wet, t_peak,
                                t_dry on a tank filled by uniform rain
                                (grid.cu:930-966, 1094-1123)
  dt                            an inactive tile's stale speed still bounds dt
                                (simulator.cu:540-547, grid.cu:836-837)
The product's host routines for the same rules (rift_map_cells,
rift_geometry) are pinned by the same closed forms.
"""
import numpy as np
import pytest

import oracle_lib as O

ND = -9999.0
EPS = float(np.float32(1.19209e-07))


def _kernel_track_windows(nbx, nby, nx, ny):
    """The cells each tile's Track block visits, enumerated as the reference
    launches it: blockDim (BLOCK_COLS, BLOCK_ROWS) = (16, 14), i = blockX *
    BLOCK_COLS + tidx + 2, j = blockY * BLOCK_COLS + tidy + 2, interior
    only (grid.cu:1224-1248)."""
    win = {}
    for by in range(nby):
        for bx in range(nbx):
            cells = set()
            for tidy in range(14):
                for tidx in range(16):
                    i, j = bx * 16 + tidx + 2, by * 16 + tidy + 2
                    if 2 <= i <= nx - 3 and 2 <= j <= ny - 3:
                        cells.add((i, j))
            win[by * nbx + bx] = cells
    return win


@pytest.mark.parametrize("kind,value", [(0, 1e-3), (1, 0.7)])
def test_track_brute_force_single_cell_maps(kind, value):
    nx, ny = 52, 74                     # 48 x 70 interior cells
    nbx, nby = (nx + 15) // 16, (ny + 13) // 14
    nt = nbx * nby
    win = _kernel_track_windows(nbx, nby, nx, ny)
    dry = np.zeros(nt, np.uint8)
    inactive = np.zeros(nt, np.uint8)
    inactive[0] = 1                     # tile 0 is never tracked (-0 == 0)
    fill = 0.0 if kind == 0 else ND
    woke_any = np.zeros((ny, nx), bool)
    for j in range(2, ny - 2):
        for i in range(2, nx - 2):
            m = np.full((ny, nx), fill)
            m[j, i] = value
            wet, _ = O.track_and_grow(nbx, nby, dry, inactive, [(kind, m)])
            want = {t for t in range(1, nt) if (i, j) in win[t]}
            got = set(np.nonzero(wet)[0].tolist())
            assert got == want, (i, j, got, want)
            woke_any[j, i] = bool(got)
    # rows 16, 17, 32, 33, ... lie in no window: the stride is 16 rows
    for j in range(2, ny - 2):
        assert woke_any[j, 2:nx - 2].any() == ((j - 2) % 16 < 14), j
    # a cell of tile row 2 by the 14-row tiling (row 30) wakes tile row 1
    m = np.full((ny, nx), fill)
    m[30, 20] = value
    wet, _ = O.track_and_grow(nbx, nby, dry, inactive, [(kind, m)])
    assert set(np.nonzero(wet)[0].tolist()) == {1 * nbx + 1}


def test_track_thresholds_and_skips():
    nx, ny = 52, 74
    nbx, nby = (nx + 15) // 16, (ny + 13) // 14
    nt = nbx * nby
    t = 2 * nbx + 1
    cell = (2 * 16 + 5, 1 * 16 + 5)        # (j, i) inside tile t's window
    inactive = np.zeros(nt, np.uint8)
    dry = np.zeros(nt, np.uint8)

    def woken(kind, v, wet=dry, act=inactive):
        m = np.full((ny, nx), 0.0 if kind == 0 else ND)
        m[cell] = v
        w, _ = O.track_and_grow(nbx, nby, wet, act, [(kind, m)])
        return bool(w[t]) and not wet[t]

    # rain wakes above epsilon only; surge above 0 and not nodata
    assert woken(0, 2 * EPS) and not woken(0, EPS) and not woken(0, 0.0)
    assert woken(1, 1e-30) and not woken(1, 0.0) and not woken(1, -0.5) and not woken(1, ND)
    # active or already-wet tiles are not tracked
    act = inactive.copy()
    act[t] = 1
    assert not woken(1, 1.0, act=act)


def _corner_rule(c, scale, allow_nodata):
    vals = [v for v in c if v != ND]
    if len(vals) > 2:
        return sum(vals) / len(vals) * scale
    return ND if allow_nodata else 0.0


@pytest.mark.parametrize("allow_nodata,scale", [(True, 1.0), (False, 1.0 / 3600.0 / 1000.0)])
def test_map_cells_nodata_patterns(rift_host, allow_nodata, scale):
    rng = np.random.default_rng(7)
    nb = 33
    corners = rng.uniform(0.5, 4.0, (nb, nb))
    corners[rng.random((nb, nb)) < 0.45] = ND
    cells = rift_host.map_cells(corners, scale, allow_nodata, ND)
    seen = set()
    for j in range(2, nb + 1):
        for i in range(2, nb + 1):
            c = (corners[j - 2, i - 2], corners[j - 2, i - 1], corners[j - 1, i - 2], corners[j - 1, i - 1])
            seen.add(tuple(v == ND for v in c))
            assert cells[j, i] == pytest.approx(_corner_rule(c, scale, allow_nodata), rel=1e-15), (i, j)
    assert len(seen) == 16                     # every nodata pattern occurred
    # ghost cells hold the fill value
    fill = ND if allow_nodata else 0.0
    assert (cells[:2] == fill).all() and (cells[:, :2] == fill).all()
    assert (cells[-2:] == fill).all() and (cells[:, -2:] == fill).all()


def test_oracle_map_read_matches_the_corner_rule():
    rng = np.random.default_rng(8)
    nb = 25
    corners = rng.uniform(0.5, 4.0, (nb, nb))
    corners[rng.random((nb, nb)) < 0.45] = ND
    o = O.Oracle(np.zeros((nb, nb)), 1.0, flags=O.SURGE_GRID)
    o.set_map_series(O.MAP_SURGE, np.stack([corners, corners]), 10, 100.0)
    s = o.field("surge")                       # factor 0 of two equal maps
    for j in range(2, nb + 1):
        for i in range(2, nb + 1):
            c = (corners[j - 2, i - 2], corners[j - 2, i - 1], corners[j - 1, i - 2], corners[j - 1, i - 1])
            assert s[j, i] == pytest.approx(_corner_rule(c, 1.0, True), rel=1e-15), (i, j)


def _flat_oracle(n, flags, **kw):
    return O.Oracle(np.zeros((n + 1, n + 1)), 1.0, flags=flags, **kw)


def test_interpolated_surge_is_linear_in_time_and_freezes_at_surge_tf():
    n, dt_map, tmax = 24, 10, 40.0
    N = n + 1
    maps = np.full((5, N, N), ND)
    for k in range(5):
        maps[k, :, :4] = 1.0 + 0.5 * k          # west strip rising 0.5 m per map
    o = _flat_oracle(n, O.SURGE_GRID | O.TILE_ACCEL)
    o.set_map_series(O.MAP_SURGE, maps, dt_map, tmax)
    seen_frozen = False
    for _ in range(4000):
        o.step()
        t, _ = o.time
        s = o.field("surge")[10, 3]
        want = 1.0 + 0.5 * (t / dt_map) if t < tmax else 3.0
        assert s == pytest.approx(want, rel=1e-12, abs=1e-12), (t, s, want)
        assert o.field("surge")[10, 10] == ND   # nodata stays nodata
        seen_frozen |= t >= tmax
        if t > tmax + 5.0:
            break
    assert seen_frozen


def test_held_rain_map_comes_from_the_map_at_t_plus_dt():
    # steps much shorter than the map interval (a tank of still water): rain
    # in [k dt, (k+1) dt) comes from map k+1, zero from surge_tf on
    n, dt_map, tmax = 16, 5, 20.0
    N = n + 1
    maps = np.stack([np.full((N, N), 10.0 * (k + 1)) for k in range(6)])   # mm/hr
    o = _flat_oracle(n, O.RAIN_GRID | O.TILE_ACCEL, h0=np.full((N, N), 1.0))
    o.set_map_series(O.MAP_RAIN, maps, dt_map, tmax)
    assert o.field("rainmap")[8, 8] == pytest.approx(20.0 / 3.6e6, rel=1e-14)   # map 1 at t = 0
    for _ in range(3000):
        o.step()
        t, dt = o.time
        assert dt < dt_map                     # one map boundary per step at most
        r = o.field("rainmap")[8, 8]
        want = 0.0 if t >= tmax else 10.0 * (int(t // dt_map) + 2) / 3.6e6
        assert r == pytest.approx(want, rel=1e-14, abs=1e-20), (t, r, want)
        if t > tmax + 2:
            break


def test_held_rain_map_advances_one_map_per_step():
    # a dry tank steps dt = dx / kappa = 10 s over 5 s maps: p_update moves
    # the series one map per call (GridSeries.cu:270-282), so it lags behind t
    n, dt_map, tmax = 16, 5, 100.0
    N = n + 1
    maps = np.stack([np.full((N, N), 10.0 * (k + 1)) for k in range(30)])
    o = _flat_oracle(n, O.RAIN_GRID | O.TILE_ACCEL)
    o.set_map_series(O.MAP_RAIN, maps, dt_map, tmax)
    idx = []
    for _ in range(5):
        o.step()
        t, dt = o.time
        idx.append((t, int(round(o.field("rainmap")[8, 8] * 3.6e6 / 10.0)) - 1))
    # first step dt = 1e-6 (no advance), then one map per 10 s step
    assert [k for _, k in idx] == [1, 2, 3, 4, 5]
    t_end, k_end = idx[-1]
    assert k_end < int(t_end // dt_map) + 1         # behind the map of t + dt


def test_surge_override_sets_depth_to_surge_minus_du_dt():
    # dry flat tank with uniform rain r: before the override dU = r in every
    # cell, so Integrate_1 leaves h = surge - r dt under the surge map and
    # h = r dt elsewhere (w* = surge + BC - dw dt, grid.cu:894-904)
    n = 32
    N = n + 1
    surge = 0.5
    maps = np.full((2, N, N), ND)
    maps[:, :, :4] = surge
    o = _flat_oracle(n, O.SURGE_GRID | O.RAIN_AVG | O.EULER)
    o.set_series_hours(O.SER_HYETO, [0.0, 10.0], [36.0, 36.0])     # 1e-5 m/s
    o.set_map_series(O.MAP_SURGE, maps, 100, 1000.0)
    o.set_time(5.0, 0.1)
    o.step()
    t, dt = o.time
    r = float(o.rates()[0])
    h = o.depth()
    assert dt > 1.0                            # a dry tank: dt = dx / kappa
    for j in range(4, n - 2):
        for i in (2, 3, 4):                    # cells with four surge corners
            assert h[j, i] == pytest.approx(surge - r * dt, rel=1e-12), (i, j)
        assert h[j, 12] == pytest.approx(r * dt, rel=1e-12)


def test_ghost_origin_tiles_keep_their_wet_flag():
    # 32 x 32 interior: nx = 36, tile column 2 starts at the ghost column 34;
    # Integrate_2's thread (0,0) of those tiles is outside the interior, so
    # their flag keeps the value InitGrid gave it (wet), and a dry domain keeps
    # them and their neighbours active
    o = _flat_oracle(32, O.TILE_ACCEL)
    assert (o.nx, o.nbx, o.nby) == (36, 3, 3)
    o.step(6)
    wet, act, _ = o.tiles()
    assert set(np.nonzero(wet)[0].tolist()) == {2, 5, 8}
    assert {int(v) for v in act if v > -1} == {0, 1, 2, 4, 5, 7, 8}


def _plane(x, y):
    return 2.0 + 0.3 * x - 0.7 * y


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_geometry_of_a_linear_dem(which, rift_host):
    b = 21
    jt, it = np.mgrid[0:b, 0:b].astype(float)
    dem = _plane(it, jt)
    if which == "oracle":
        o = O.Oracle(dem, 1.0)
        BX, BY = o.faces()
        BC = o.field("bc")
        nx, ny = o.nx, o.ny
    else:
        geo = rift_host.geometry(dem)
        BX, BY, nx, ny = geo["BX"], geo["BY"], geo["nx"], geo["ny"]
        BC = None
    for j in range(2, ny - 2):
        # west face of cell i at column i: the plane at (i - 2, j - 1.5)
        for i in range(2, nx - 1):       # faces 2 .. nx-2
            assert BX[j, i] == pytest.approx(_plane(i - 2, j - 1.5), abs=1e-13), (i, j)
        # ghost faces repeat the nearest interior faces (simulator.cu:404-414)
        assert BX[j, 1] == BX[j, 3] and BX[j, 0] == BX[j, 2] and BX[j, nx - 1] == BX[j, nx - 3]
    for i in range(2, nx - 2):
        for j in range(2, ny - 1):
            assert BY[j, i] == pytest.approx(_plane(i - 1.5, j - 2), abs=1e-13), (i, j)
        assert BY[1, i] == BY[3, i] and BY[0, i] == BY[2, i] and BY[ny - 1, i] == BY[ny - 3, i]
    if BC is not None:
        # the cell bed is the plane at the cell centre
        for j in range(2, ny - 2):
            for i in range(2, nx - 2):
                assert BC[j, i] == pytest.approx(_plane(i - 1.5, j - 1.5), abs=1e-13)


@pytest.mark.parametrize("euler", [True, False], ids=["euler", "heun"])
def test_summary_maps_of_a_tank_filled_by_rain(euler):
    # a flat tank under uniform rain rises monotonically at the centre: h_max
    # is the last depth, t_wet the time of the first depth above epsilon, and
    # in Euler mode t_peak the last step's time and t_dry the last time the
    # depth exceeded 0.0762 m (grid.cu:930-966, 1094-1123)
    # (256 m across: the outflow boundaries' drawdown does not reach the
    # centre within the 100 s)
    flags = O.RAIN_AVG | O.SAVE_MAX | O.SAVE_ARRIVAL | (O.EULER if euler else 0)
    o = O.Oracle(np.zeros((65, 65)), 4.0, flags=flags)
    o.set_series_hours(O.SER_HYETO, [0.0, 10.0], [3600.0, 3600.0])   # 1e-3 m/s
    c = (34, 34)
    hist = []
    while not hist or hist[-1][0] < 100.0:
        o.step()
        hist.append((o.time[0], o.depth()[c]))
    assert all(b[1] > a[1] for a, b in zip(hist[1:], hist[2:]))     # monotone rise
    h_last = hist[-1][1]
    assert h_last > 0.0762
    # fmaxf on doubles (grid.cu:941, 1101): the maximum is held rounded to float
    assert o.field("h_max")[c] == float(np.float32(h_last))
    t_first_wet = next(t for t, h in hist if h > EPS)
    assert o.field("t_wet")[c] == pytest.approx(t_first_wet, rel=1e-12)
    assert abs(o.field("q_max")[c]) < 1e-9           # still water at the centre
    if euler:
        assert o.field("t_peak")[c] == pytest.approx(hist[-1][0], rel=1e-12)
        assert o.field("t_dry")[c] == pytest.approx(max(t for t, h in hist if h > 0.0762), rel=1e-12)


def test_stale_speed_of_an_inactive_tile_bounds_dt():
    # deep water in one tile gives it the largest speed; the tile is then made
    # inactive and the deep water removed: the next flux pass updates only the
    # active tiles, and ComputeTimestep's max over every tile still sees the
    # stale speed (simulator.cu:540-547)
    n = 64
    N = n + 1
    h0 = np.full((N, N), 0.1)
    h0[16:30, 18:34] = 4.0
    o = O.Oracle(np.zeros((N, N)), 1.0, h0=h0, flags=O.TILE_ACCEL, t0=2.0)
    o.phase("grow")
    o.phase("compute_fluxes")
    mx, my = o.tile_speeds()
    deep = int(np.argmax(mx))
    stale = max(mx[deep], my[deep])
    wet, act, _ = o.tiles()
    aval = np.arange(o.ntiles)
    aval[deep] = -deep
    o.set_tiles(wet, aval)
    w = o.field("w")
    w[w > 1.0] = 0.1
    o.set_field("w", w)
    o.phase("compute_fluxes")
    mx2, my2 = o.tile_speeds()
    others = max(np.delete(mx2, deep).max(), np.delete(my2, deep).max())
    assert mx2[deep] == mx[deep] and others < 0.5 * stale
    o.phase("compute_timestep")
    t, dt = o.time
    kap = O.lib().orc_kappa(1.0, 1.0)
    c = lambda v: 1.0 / float(np.float32(max(np.float32(v / 0.25), np.float32(kap))))
    assert dt == min(c(mx2.max()), c(my2.max()))
    assert dt < 0.6 * c(others)                      # the stale speed decides


@pytest.fixture(scope="module")
def rift_host():
    """The product library's host routines (no device work)."""
    from rift_amd import kp
    kp.load()
    return kp
