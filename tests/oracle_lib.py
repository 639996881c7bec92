"""ctypes binding of the CPU oracle (oracle/kp_oracle.{h,c}).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product never loads the oracle.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

FAITHFUL, TEXTBOOK = 0, 1

HYDROGRAPH = 1 << 0
RAIN_AVG = 1 << 1
RAIN_GRID = 1 << 2
DRAIN_AVG = 1 << 3
DRAIN_GRID = 1 << 4
INFILTRATION = 1 << 5
SURGE_GRID = 1 << 6
EULER = 1 << 7
SAVE_MAX = 1 << 8
SAVE_ARRIVAL = 1 << 9
TILE_ACCEL = 1 << 10
H_INIT = 1 << 11
RAIN_INTERP = 1 << 12

FIELDS = ["w", "hu", "hv", "depth", "bc", "n", "K", "F", "dw", "dhu", "dhv", "G", "dF",
          "h_max", "q_max", "t_wet", "t_peak", "t_dry", "surge", "rainmap", "drainmap", "q"]
SER_HYETO, SER_DRAIN, SER_HYDRO = 0, 1, 2
MAP_RAIN, MAP_DRAIN, MAP_SURGE = 0, 1, 2


class OrcParams(C.Structure):
    _fields_ = [("b_nx", C.c_int), ("b_ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("kappa", C.c_double), ("courant", C.c_double), ("psi", C.c_double),
                ("dtheta", C.c_double), ("t0", C.c_double), ("nodata", C.c_double),
                ("n_const", C.c_double), ("flags", C.c_uint), ("mode", C.c_int)]


_LIBS = {}
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_ubyte)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(fp32=False):
    key = "32" if fp32 else "64"
    if key in _LIBS:
        return _LIBS[key]
    path = os.path.join(ORACLE_DIR, "liboracle32.so" if fp32 else "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.orc_signf.restype = C.c_double
    L.orc_signf.argtypes = [C.c_double]
    L.orc_minmod.restype = C.c_double
    L.orc_minmod.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_limiter.restype = C.c_double
    L.orc_limiter.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_velocity.restype = C.c_double
    L.orc_velocity.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_kp_flux.restype = C.c_double
    L.orc_kp_flux.argtypes = [C.c_double] * 6
    L.orc_construct.argtypes = [C.c_int, _D, C.c_double, C.c_double, _D]
    L.orc_face.argtypes = [C.c_int, C.c_int, _D, _D, C.c_double, C.c_double, _D]
    L.orc_kappa.restype = C.c_double
    L.orc_kappa.argtypes = [C.c_double, C.c_double]
    L.orc_grow_tiles.argtypes = [C.c_int, C.c_int, _U8, _I]
    L.orc_track_tiles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _D, C.c_int, C.c_double, _I, _U8]
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [C.POINTER(OrcParams), _D, _D, _D, _D]
    L.orc_destroy.argtypes = [C.c_void_p]
    for name in ("orc_nx", "orc_ny", "orc_ntiles", "orc_nbx"):
        getattr(L, name).restype = C.c_int
        getattr(L, name).argtypes = [C.c_void_p]
    L.orc_set_series_hours.argtypes = [C.c_void_p, C.c_int, _F, _F, C.c_int]
    L.orc_set_sources.argtypes = [C.c_void_p, _I, C.c_long]
    L.orc_set_map_series.argtypes = [C.c_void_p, C.c_int, _D, C.c_int, C.c_int, C.c_double]
    L.orc_set_map_series_range.argtypes = [C.c_void_p, C.c_int, C.c_int, _D, C.c_int, C.c_int,
                                           C.c_double]
    for name in ("orc_step", "orc_update_source", "orc_grow", "orc_compute_fluxes",
                 "orc_compute_timestep", "orc_integrate_1", "orc_integrate_2",
                 "orc_apply_boundaries", "orc_update_maps"):
        getattr(L, name).argtypes = [C.c_void_p]
    L.orc_run.argtypes = [C.c_void_p, C.c_int]
    L.orc_get_time.argtypes = [C.c_void_p, _D, _D]
    L.orc_set_time.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.orc_get_field.argtypes = [C.c_void_p, C.c_int, _D]
    L.orc_set_field.argtypes = [C.c_void_p, C.c_int, _D]
    L.orc_get_faces.argtypes = [C.c_void_p, _D, _D]
    L.orc_get_tiles.argtypes = [C.c_void_p, _U8, _I, _I]
    L.orc_set_tiles.argtypes = [C.c_void_p, _U8, _I]
    L.orc_get_tile_speeds.argtypes = [C.c_void_p, _D, _D]
    L.orc_get_rates.argtypes = [C.c_void_p, _F]
    L.orc_mass.argtypes = [C.c_void_p, _D]
    L.orc_budget.argtypes = [C.c_void_p, _D]
    L.orc_budget_reset.argtypes = [C.c_void_p]
    _LIBS[key] = L
    return L


def dptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


class Oracle:
    """One oracle simulation.  Arrays use the reference's orientation:
    row 0 = south; corner rasters are (b_ny, b_nx), internal fields (ny, nx)."""

    def __init__(self, dem, dx, *, h0=None, n=None, n_const=0.0, K=None, psi=0.0,
                 dtheta=0.0, courant=0.25, t0=0.0, flags=TILE_ACCEL, mode=FAITHFUL,
                 fp32=False, nodata=-9999.0, kappa=None, dy=None):
        self.L = lib(fp32)
        dem = np.ascontiguousarray(dem, dtype=np.float64)
        self.b_ny, self.b_nx = dem.shape
        dy = dx if dy is None else dy
        if h0 is not None:
            flags |= H_INIT
        if K is not None:
            flags |= INFILTRATION
        p = OrcParams(self.b_nx, self.b_ny, dx, dy,
                      self.L.orc_kappa(dx, dy) if kappa is None else kappa,
                      courant, psi, dtheta, t0, nodata, n_const, flags, mode)
        self.params = p
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                      for a in (h0, n, K)]
        h0a, na, Ka = self._keep
        self.s = self.L.orc_create(C.byref(p), dptr(dem), dptr(h0a), dptr(na), dptr(Ka))
        self.nx = self.L.orc_nx(self.s)
        self.ny = self.L.orc_ny(self.s)
        self.ntiles = self.L.orc_ntiles(self.s)
        self.nbx = self.L.orc_nbx(self.s)
        self.nby = self.ntiles // self.nbx

    def __del__(self):
        if getattr(self, "s", None):
            self.L.orc_destroy(self.s)
            self.s = None

    def set_series_hours(self, which, hours, values):
        h = np.ascontiguousarray(hours, dtype=np.float32)
        v = np.ascontiguousarray(values, dtype=np.float32)
        assert self.L.orc_set_series_hours(self.s, which, h.ctypes.data_as(_F),
                                           v.ctypes.data_as(_F), len(h)) == 0

    def set_sources(self, idx):
        a = np.ascontiguousarray(idx, dtype=np.int32)
        self.L.orc_set_sources(self.s, a.ctypes.data_as(_I), len(a))

    def set_map_series(self, which, maps, deltat, tmax, first=0):
        """Corner maps of file indices first, first+1, ...; the first
        p_update runs at the oracle's current time."""
        m = np.ascontiguousarray(maps, dtype=np.float64)
        assert self.L.orc_set_map_series_range(self.s, which, int(first), dptr(m), m.shape[0],
                                               int(deltat), float(tmax)) == 0

    def step(self, n=1):
        self.L.orc_run(self.s, n)

    def phase(self, name):
        getattr(self.L, "orc_" + name)(self.s)

    @property
    def time(self):
        t, dt = C.c_double(), C.c_double()
        self.L.orc_get_time(self.s, C.byref(t), C.byref(dt))
        return t.value, dt.value

    def set_time(self, t, dt):
        self.L.orc_set_time(self.s, t, dt)

    def field(self, name):
        out = np.empty((self.ny, self.nx), dtype=np.float64)
        assert self.L.orc_get_field(self.s, FIELDS.index(name), dptr(out)) == 0
        return out

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert self.L.orc_set_field(self.s, FIELDS.index(name), dptr(a)) == 0

    def faces(self):
        bx = np.empty((self.ny, self.nx + 1))
        by = np.empty((self.ny + 1, self.nx))
        self.L.orc_get_faces(self.s, dptr(bx), dptr(by))
        return bx, by

    def tiles(self):
        wet = np.empty(self.ntiles, dtype=np.uint8)
        act = np.empty(self.ntiles, dtype=np.int32)
        n = C.c_int()
        self.L.orc_get_tiles(self.s, wet.ctypes.data_as(_U8), act.ctypes.data_as(_I), C.byref(n))
        return wet, act, n.value

    def set_tiles(self, wet, act):
        w = np.ascontiguousarray(wet, dtype=np.uint8)
        a = np.ascontiguousarray(act, dtype=np.int32)
        self.L.orc_set_tiles(self.s, w.ctypes.data_as(_U8), a.ctypes.data_as(_I))

    def active_set(self):
        wet, act, n = self.tiles()
        return set(int(v) for v in act if v > -1)

    def tile_speeds(self):
        mx = np.empty(self.ntiles)
        my = np.empty(self.ntiles)
        self.L.orc_get_tile_speeds(self.s, dptr(mx), dptr(my))
        return mx, my

    def rates(self):
        out = np.empty(3, dtype=np.float32)
        self.L.orc_get_rates(self.s, out.ctypes.data_as(_F))
        return out

    def mass(self):
        out = np.empty(4)
        self.L.orc_mass(self.s, dptr(out))
        return out

    def budget(self):
        """[stored, infiltrated, total at reset, rain, source, drain, outflow,
        exchange, residual] in m^3."""
        out = np.empty(9)
        self.L.orc_budget(self.s, dptr(out))
        return out

    def budget_reset(self):
        self.L.orc_budget_reset(self.s)

    def depth(self):
        return self.field("depth")

    def interior(self, a):
        return a[2:self.ny - 2, 2:self.nx - 2]


def track_and_grow(nbx, nby, wet, act, maps, nodata=-9999.0, fp32=False):
    """The oracle's Track (grid.cu:1224-1276) and Grow (grid.cu:1188-1322) on
    given tile flags: wet (ntiles u8) and act (ntiles: 1 active, 0 not) of the
    last Grow, maps = [(kind, (ny, nx) map)] with kind 0 rain, 1 surge.
    Returns (wet after Track, active flags after Grow)."""
    L = lib(fp32)
    nt = nbx * nby
    w = np.ascontiguousarray(wet, dtype=np.uint8).copy()
    aval = np.where(np.asarray(act) > 0, np.arange(nt), -np.arange(nt)).astype(np.int32)
    for kind, m in maps:
        a = np.ascontiguousarray(m, dtype=np.float64)
        ny, nx = a.shape
        L.orc_track_tiles(nx, ny, nbx, nby, dptr(a), int(kind), float(nodata), aval.ctypes.data_as(_I),
                          w.ctypes.data_as(_U8))
    out = np.empty(nt, dtype=np.int32)
    L.orc_grow_tiles(nbx, nby, w.ctypes.data_as(_U8), out.ctypes.data_as(_I))
    return w, (out >= 0).astype(np.uint8) | (np.arange(nt) == 0)
