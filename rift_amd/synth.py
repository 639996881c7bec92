"""Seeded synthetic inputs for the benchmark workloads.

The reference reads every input from files (io.cc, source.cc); these
generators produce the same kinds of inputs -- corner rasters with row 0 =
south, CSV-style (hours, value) series -- from a counter-based SplitMix64
stream, so that a seed names one exact input on every machine.

Workloads (all rasters are (N+1) x (N+1) DEM points, so the interior is
exactly N x N cells):
  dam_break       256 x 256, dx = 1 m, flat bed, h0 = 1.0 m west of x = 128,
                  0.1 m east, n = 0, tf = 20 s
  rain_basin      N x N (2048 for the benchmark), dx = 10 m, diamond-square
                  DEM (roughness 0.6) scaled to 0-60 m with a 3 m deep channel,
                  n = 0.035, dry start, triangular hyetograph 0 -> 80 -> 0 mm/hr
                  over 1 h, Green-Ampt K lognormal (median 5 mm/hr, sigma_ln
                  0.8), psi = 0.11 m, dtheta = 0.3, tf = 2 h
  coastal_surge   N x N (8192), dx = 5 m, ramp -5 -> +30 m plus fractal noise
                  (amplitude 4 m), n = 0.02 below 0 m and 0.06 above, surge
                  depth maps every 600 s on the low edge, one rectangular
                  hydrograph source 0 -> 500 cfs, tf = 6 h
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n, stream=0):
    """n 64-bit outputs of SplitMix64 at counters 0..n-1 of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, stream=0):
    """Uniform doubles in (0, 1)."""
    z = splitmix64(seed, n, stream)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed, n, stream=0):
    """Standard normals by Box-Muller on two uniform streams."""
    u1 = uniform(seed, n, 2 * stream + 1)
    u2 = uniform(seed, n, 2 * stream + 2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def diamond_square(levels, roughness, seed):
    """(2^levels + 1)^2 fractal surface in [0, 1] by diamond-square."""
    N = (1 << levels) + 1
    a = np.zeros((N, N))
    a[0, 0], a[0, -1], a[-1, 0], a[-1, -1] = uniform(seed, 4, 0) - 0.5
    step, scale, lvl = N - 1, 1.0, 1
    while step > 1:
        half = step // 2
        ys = np.arange(half, N, step)
        Y, X = np.meshgrid(ys, ys, indexing="ij")
        avg = 0.25 * (a[Y - half, X - half] + a[Y - half, X + half] + a[Y + half, X - half] +
                      a[Y + half, X + half])
        a[Y, X] = avg + scale * (uniform(seed, Y.size, 2 * lvl).reshape(Y.shape) - 0.5)
        for y0, x0 in ((0, half), (half, 0)):
            Y, X = np.meshgrid(np.arange(y0, N, step), np.arange(x0, N, step), indexing="ij")
            acc = np.zeros(Y.shape)
            cnt = np.zeros(Y.shape)
            for dy, dx in ((-half, 0), (half, 0), (0, -half), (0, half)):
                yy, xx = Y + dy, X + dx
                ok = (yy >= 0) & (yy < N) & (xx >= 0) & (xx < N)
                acc[ok] += a[yy[ok], xx[ok]]
                cnt[ok] += 1
            stream = 2 * lvl + 1 + (0 if y0 == 0 else 1000)
            a[Y, X] = acc / cnt + scale * (uniform(seed, Y.size, stream).reshape(Y.shape) - 0.5)
        scale *= 2.0 ** (-roughness)
        step = half
        lvl += 1
    a -= a.min()
    a /= max(a.max(), 1e-300)
    return a


def fractal_dem(n_cells, roughness, seed):
    """(n_cells+1)^2 fractal surface in [0,1]; n_cells must be a power of two
    or the surface is cropped from the next power of two."""
    levels = int(np.ceil(np.log2(n_cells)))
    a = diamond_square(levels, roughness, seed)
    a = a[:n_cells + 1, :n_cells + 1]
    a -= a.min()
    return a / max(a.max(), 1e-300)


def dam_break(n_cells=256):
    N = n_cells + 1
    dem = np.zeros((N, N))
    h0 = np.where(np.arange(N)[None, :] < n_cells // 2, 1.0, 0.1) * np.ones((N, 1))
    return {"dem": dem, "h0": h0, "dx": 1.0, "n_const": 0.0, "tf": 20.0}


def carved_channel(dem, depth, seed):
    """Subtract `depth` along a meandering west-east channel."""
    N = dem.shape[0]
    x = np.arange(N)
    phase = 2 * np.pi * uniform(seed, 1, 7)[0]
    yc = 0.5 * (N - 1) + 0.12 * (N - 1) * np.sin(2 * np.pi * 1.5 * x / (N - 1) + phase)
    half_w = max(1.5, N / 256.0)
    Y = np.arange(N)[:, None]
    mask = np.abs(Y - yc[None, :]) <= half_w
    out = dem.copy()
    out[mask] -= depth
    return out


def rain_basin(n_cells=2048, seed=2):
    """Rain on a fractal basin with Green-Ampt infiltration (dx = 10 m)."""
    dem = 60.0 * fractal_dem(n_cells, 0.6, seed)
    dem = carved_channel(dem, 3.0, seed)
    N = n_cells + 1
    K = 5.0 * np.exp(0.8 * normal(seed * 10, N * N).reshape(N, N))
    hyeto_hours = np.array([0.0, 0.5, 1.0], dtype=np.float32)
    hyeto_mmhr = np.array([0.0, 80.0, 0.0], dtype=np.float32)
    return {"dem": dem, "K": K, "dx": 10.0, "n_const": 0.035, "psi": 0.11, "dtheta": 0.3,
            "hyeto_hours": hyeto_hours, "hyeto_mmhr": hyeto_mmhr, "tf": 7200.0}


SURGE_NODATA = -9999.0


def surge_strip_rows(n_cells):
    """Corner rows of the surge strip on the low (south) edge."""
    return max(4, (n_cells + 1) // 64)


def surge_depth(k, surge_dt=600):
    """Surge depth of map k: 2.5 sin(pi t / 2 h) at t = k surge_dt, clipped >= 0."""
    return max(0.0, 2.5 * np.sin(np.pi * (k * float(surge_dt)) / (2 * 3600.0)))


def surge_map(n_cells, k, surge_dt=600):
    """Corner map k of the coastal surge series: the surge depth on the strip
    of the low edge, nodata elsewhere ((n_cells+1)^2 doubles)."""
    N = n_cells + 1
    m = np.full((N, N), SURGE_NODATA)
    m[:surge_strip_rows(n_cells), :] = surge_depth(k, surge_dt)
    return m


def coastal_surge(n_cells=8192, seed=3, nmaps=37, with_maps=True):
    """Coastal ramp with surge maps on the low (south) edge and a hydrograph.
    with_maps=False leaves the maps out (`surge_map` makes map k on demand:
    37 maps at 8192 x 8192 would be ~20 GB of doubles)."""
    N = n_cells + 1
    y = np.linspace(0.0, 1.0, N)[:, None]
    dem = -5.0 + 35.0 * y + 4.0 * (fractal_dem(n_cells, 0.6, seed) - 0.5)
    n_map = np.where(dem < 0.0, 0.02, 0.06)
    hydro_hours = np.array([0.0, 1.0, 3.0, 6.0], dtype=np.float32)
    hydro_cfs = np.array([0.0, 500.0, 500.0, 0.0], dtype=np.float32)
    out = {"dem": dem, "n": n_map, "dx": 5.0, "surge_dt": 600,
           "surge_tf": 21600.0, "hydro_hours": hydro_hours, "hydro_cfs": hydro_cfs,
           "source_rect": (N // 2 - 10, N // 2 + 10, N // 2, N // 2 + 20), "tf": 21600.0}
    if with_maps:
        out["surge_maps"] = np.stack([surge_map(n_cells, k) for k in range(nmaps)])
    return out


def source_cells(nx, rect):
    """Internal-grid cell indices j*nx + i of a source rectangle
    (i0, i1, j0, j1), i outer and j inner (InitializeSources order)."""
    i0, i1, j0, j1 = rect
    return np.array([j * nx + i for i in range(i0, i1 + 1) for j in range(j0, j1 + 1)], dtype=np.int32)


def series_hours_to_seconds(hours, values):
    """Source::ReadSource's conversion (source.cc:36-63): hours*60*60 in float
    and a trailing zero-rate point when the last rate is nonzero."""
    t = np.asarray(hours, dtype=np.float32) * np.float32(60) * np.float32(60)
    v = np.asarray(values, dtype=np.float32)
    if len(v) and v[-1] != np.float32(0):
        est = np.float32(0)
        for k in range(1, len(t)):
            est = np.float32(est + np.float32(t[k] - t[k - 1]))
        est = np.float32(est / np.float32(np.float32(len(t)) - np.float32(1)))
        t = np.append(t, np.float32(t[-1] + est))
        v = np.append(v, np.float32(0))
    return t.astype(np.float32), v.astype(np.float32)


def _write_asc(path, a, cellsize, nodata=-9999.0):
    """ESRI-ASCII raster of a corner array (row 0 = south), lower-left corner
    at the origin; values in 17 significant digits so they read back exactly."""
    a = np.asarray(a, dtype=np.float64)
    nr, nc = a.shape
    with open(path, "w") as f:
        f.write(f"ncols {nc}\nnrows {nr}\nxllcorner 0\nyllcorner 0\n"
                f"cellsize {cellsize!r}\nNODATA_value {nodata!r}\n")
        for j in range(nr - 1, -1, -1):
            f.write(" ".join("%.17g" % v for v in a[j]) + "\n")


def _write_csv(path, hours, values, header):
    with open(path, "w") as f:
        f.write(f"# {header}\n")
        for h, v in zip(hours, values):
            f.write("%.9g,%.9g\n" % (float(h), float(v)))


def write_case(case, outdir, name="case", **keywords):
    """Writes a generated workload as the files a RIFT run reads: the DEM and
    corner rasters (`.asc`), CSV series, surge maps `surge-###.txt`, and the
    keyword file `<name>.cfg` (projected metres, lower-left corner at the
    origin).  `keywords` add or override configuration keywords (e.g.
    save_max=1, dt_print=60).  Returns the path of the keyword file."""
    import os
    os.makedirs(outdir, exist_ok=True)
    dx = float(case["dx"])
    cfg = {"DEM": "dem.asc", "projected": 1, "t0": 0, "tf": case["tf"], "output": "."}
    _write_asc(os.path.join(outdir, "dem.asc"), case["dem"], dx)
    if "n" in case:
        _write_asc(os.path.join(outdir, "n.asc"), case["n"], dx)
        cfg["n"] = "n.asc"
    else:
        cfg["n"] = case.get("n_const", 0.0)
    if "h0" in case:
        _write_asc(os.path.join(outdir, "h0.asc"), case["h0"], dx)
        cfg["h0"] = "h0.asc"
    if "K" in case:
        _write_asc(os.path.join(outdir, "K.asc"), case["K"], dx)
        cfg.update({"K": "K.asc", "psi": case["psi"], "dtheta": case["dtheta"]})
    if "hyeto_hours" in case:
        _write_csv(os.path.join(outdir, "hyetograph.csv"), case["hyeto_hours"], case["hyeto_mmhr"],
                   "hours, mm/hr")
        cfg["hyetograph"] = "hyetograph.csv"
    if "hydro_hours" in case:
        _write_csv(os.path.join(outdir, "hydrograph.csv"), case["hydro_hours"], case["hydro_cfs"],
                   "hours, cfs")
        i0, i1, j0, j1 = case["source_rect"]
        # Coord2Index: i = (int)((X - h_xll) / cellsize + 0.5), h_xll = dx / 2
        x = lambda i: 0.5 * dx + i * dx
        cfg["hydrograph"] = "hydrograph.csv"
        cfg["source_rect_X"] = f"({x(i0)!r},{x(i1)!r})"
        cfg["source_rect_Y"] = f"({x(j0)!r},{x(j1)!r})"
    if "surge_maps" in case:
        for k, m in enumerate(case["surge_maps"]):
            _write_asc(os.path.join(outdir, "surge-%03d.txt" % k), m, dx)
        cfg.update({"surge_prefix": "surge", "surge_dt": case["surge_dt"], "surge_tf": case["surge_tf"]})
    cfg.update(keywords)
    path = os.path.join(outdir, f"{name}.cfg")
    with open(path, "w") as f:
        f.write(f"; generated workload {name}\n")
        for k, v in cfg.items():
            f.write(f"{k} = {v!r}\n" if isinstance(v, float) else f"{k} = {v}\n")
    return path


ENSEMBLE_SIDES = (512, 768, 1024, 1536, 2048)


def ensemble_scenarios(n=64, seed=4, sides=ENSEMBLE_SIDES, tf=3600.0):
    """The batched-ensemble workload: n scenarios with the side drawn
    uniformly from `sides` (dx = 10 m), a rain_basin-style fractal DEM
    (0-60 m, 3 m channel) per scenario, Manning n ~ U(0.02, 0.08), Green-Ampt
    K lognormal (median 5 mm/hr, sigma_ln 0.8), psi 0.11, dtheta 0.3, and a
    triangular hyetograph over 1 h whose peak is lognormal (median 50 mm/hr,
    sigma_ln 0.5).  Returns member dicts (the Ensemble / grid_from_rasters
    arguments plus `hours`, `mmhr`, `tf`)."""
    u = uniform(seed, 3 * n, 11)
    g = normal(seed, n, 12)
    out = []
    for k in range(n):
        side = sides[min(int(u[k] * len(sides)), len(sides) - 1)]
        dem = 60.0 * fractal_dem(side, 0.6, seed * 1000 + k)
        dem = carved_channel(dem, 3.0, seed * 1000 + k)
        N = side + 1
        K = 5.0 * np.exp(0.8 * normal(seed * 1000 + k, N * N, 13).reshape(N, N))
        n_man = 0.02 + 0.06 * u[n + k]
        peak = 50.0 * np.exp(0.5 * g[k])
        out.append({"dem": dem, "dx": 10.0, "n_const": float(n_man), "K": K, "psi": 0.11,
                    "dtheta": 0.3, "hours": np.array([0.0, 0.5, 1.0], np.float32),
                    "mmhr": np.array([0.0, peak, 0.0], np.float32), "tf": tf, "side": side})
    return out
