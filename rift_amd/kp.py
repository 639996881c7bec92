"""ctypes binding of librift_kp.so (include/rift_kp.h, include/rift_host.h).

`Grid` mirrors the reference's Grid host API (grid.h:6-69) over one handle:
construction = SetDeviceConstants + AllocateGrid, `upload_static` + `init` =
the uploads of InitSimulation + InitGrid + ApplyBoundaries, `step` =
RunSimulation's body, and the phase methods map one-to-one onto
Grow / ComputeFluxes / ComputeTimestep / Integrate_1 / Integrate_2 /
ApplyBoundaries.  There is no CPU fallback: without the built library every
call raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librift_kp.so")

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
TEXTBOOK = 1 << 16

FIELDS = ["w", "hu", "hv", "depth", "bc", "n", "K", "F", "dw", "dhu", "dhv", "G", "dF",
          "h_max", "q_max", "t_wet", "t_peak", "t_dry", "surge", "rainmap", "drainmap", "q"]
SER_HYETO, SER_DRAIN, SER_HYDRO = 0, 1, 2
MAP_RAIN, MAP_DRAIN, MAP_SURGE = 0, 1, 2
MASS_BUDGET = 1 << 13
OPT_STEP_PATH, OPT_GRAPHS, OPT_VARIANT, OPT_MASS_BUDGET, OPT_DT_KERNEL = 1, 2, 3, 4, 5
NVARIANTS = 64
KERNEL_CLASSES = ["flux1", "stage2_flux", "integrate", "boundaries", "grow", "timestep"]

# (name, restype, argtypes) of every exported entry point
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_U8 = C.POINTER(C.c_ubyte)
_VP = C.c_void_p


class RiftDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("kappa", C.c_double), ("courant", C.c_double), ("psi", C.c_double),
                ("dtheta", C.c_double), ("nodata", C.c_double), ("flags", C.c_uint),
                ("device", C.c_int)]


class UNetDesc(C.Structure):
    _fields_ = [("height", C.c_int), ("width", C.c_int), ("in_channels", C.c_int),
                ("out_channels", C.c_int), ("base_channels", C.c_int), ("levels", C.c_int),
                ("max_batch", C.c_int), ("seed", C.c_ulonglong), ("device", C.c_int)]


SIGNATURES = [
    ("rift_last_error", C.c_char_p, []),
    ("rift_create", C.c_int, [C.POINTER(RiftDesc), C.POINTER(_VP)]),
    ("rift_destroy", C.c_int, [_VP]),
    ("rift_upload_static", C.c_int, [_VP, _D, _D, _D, _D, _D]),
    ("rift_init", C.c_int, [_VP, C.c_double]),
    ("rift_set_series", C.c_int, [_VP, C.c_int, _F, _F, C.c_int]),
    ("rift_set_source_cells", C.c_int, [_VP, _I, C.c_long]),
    ("rift_set_source_value", C.c_int, [_VP, C.c_double]),
    ("rift_set_map", C.c_int, [_VP, C.c_int, C.c_int, _D]),
    ("rift_interp_map", C.c_int, [_VP, C.c_int, C.c_double]),
    ("rift_copy_map", C.c_int, [_VP, C.c_int, C.c_int, C.c_int]),
    ("rift_map_series_create", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_double]),
    ("rift_map_series_load", C.c_int, [_VP, C.c_int, C.c_int, _D]),
    ("rift_map_series_start", C.c_int, [_VP, C.c_int, C.c_double]),
    ("rift_map_series_state", C.c_int, [_VP, C.c_int, _I, _I, _D, _I]),
    ("rift_map_cells", C.c_int, [C.c_int, C.c_int, _D, C.c_double, C.c_int, C.c_double, _D]),
    ("rift_step", C.c_int, [_VP, C.c_int]),
    ("rift_update_source", C.c_int, [_VP]),
    ("rift_grow", C.c_int, [_VP]),
    ("rift_compute_fluxes", C.c_int, [_VP]),
    ("rift_compute_timestep", C.c_int, [_VP]),
    ("rift_integrate_1", C.c_int, [_VP]),
    ("rift_integrate_2", C.c_int, [_VP]),
    ("rift_apply_boundaries", C.c_int, [_VP]),
    ("rift_set_option", C.c_int, [_VP, C.c_int, C.c_int]),
    ("rift_autotune", C.c_int, [_VP, C.c_int, _I, _D]),
    ("rift_sync", C.c_int, [_VP]),
    ("rift_get_time", C.c_int, [_VP, _D, _D]),
    ("rift_set_time", C.c_int, [_VP, C.c_double, C.c_double]),
    ("rift_download", C.c_int, [_VP, C.c_int, _D]),
    ("rift_upload", C.c_int, [_VP, C.c_int, _D]),
    ("rift_get_tiles", C.c_int, [_VP, _U8, _U8, _I]),
    ("rift_set_tiles", C.c_int, [_VP, _U8, _U8]),
    ("rift_get_tile_speeds", C.c_int, [_VP, _F, _F]),
    ("rift_get_tile_dims", C.c_int, [_VP, _I, _I]),
    ("rift_get_desc", C.c_int, [_VP, C.POINTER(RiftDesc)]),
    ("rift_get_rates", C.c_int, [_VP, _F]),
    ("rift_sum_interior", C.c_int, [_VP, C.c_int, _D]),
    ("rift_sum_interior_noalloc", C.c_int, [_VP, C.c_int, _D]),
    ("rift_mass_balance", C.c_int, [_VP, _D]),
    ("rift_budget", C.c_int, [_VP, _D]),
    ("rift_budget_reset", C.c_int, [_VP]),
    ("rift_get_stream", C.c_int, [_VP, C.POINTER(_VP)]),
    ("rift_profile_enable", C.c_int, [_VP, C.c_int]),
    ("rift_kernel_stats", C.c_int, [_VP, C.c_int, C.POINTER(C.c_long), _D]),
    ("rift_wet_cell_updates", C.c_int, [_VP, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("rift_kappa", C.c_double, [C.c_double, C.c_double]),
    ("rift_geometry", C.c_int, [C.c_int, C.c_int, _D, _D, _D, C.c_double, _D, _D, _D, _D, _D, _D]),
    ("rift_corners_from_faces", C.c_int, [C.c_int, C.c_int, _D, _D, _D]),
    ("rift_read_series", C.c_int, [C.c_char_p, _F, _F, C.c_int, _I]),
    ("rift_read_raster", C.c_int, [C.c_char_p, C.c_int, _I, _I, _D, _D, C.c_long]),
    ("rift_write_grid", C.c_int, [C.c_char_p, _D, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]),
    ("rift_ensemble_create", C.c_int, [C.c_int, C.POINTER(RiftDesc), C.POINTER(_VP)]),
    ("rift_ensemble_destroy", C.c_int, [_VP]),
    ("rift_ensemble_size", C.c_int, [_VP]),
    ("rift_ensemble_member", _VP, [_VP, C.c_int]),
    ("rift_ensemble_set_tf", C.c_int, [_VP, C.c_int, C.c_double]),
    ("rift_ensemble_step", C.c_int, [_VP, C.c_int]),
    ("rift_ensemble_finished", C.c_int, [_VP, _I]),
    ("rift_ensemble_run", C.c_int, [_VP, C.c_int, C.c_long, C.POINTER(C.c_long)]),
    ("rift_config_open", C.c_int, [C.c_char_p, C.POINTER(_VP)]),
    ("rift_config_close", None, [_VP]),
    ("rift_config_has", C.c_int, [_VP, C.c_char_p]),
    ("rift_config_is_number", C.c_int, [_VP, C.c_char_p, _I]),
    ("rift_config_string", C.c_int, [_VP, C.c_char_p, C.c_char_p, C.c_int, _I]),
    ("rift_config_double", C.c_int, [_VP, C.c_char_p, _D]),
    ("rift_config_bool", C.c_int, [_VP, C.c_char_p, _I]),
    ("rift_config_complex", C.c_int, [_VP, C.c_char_p, _D, _D]),
    ("rift_sim_open", C.c_int, [C.c_char_p, C.c_double, C.POINTER(_VP)]),
    ("rift_sim_step", C.c_int, [_VP, _D]),
    ("rift_sim_run", C.c_int, [_VP, C.c_double]),
    ("rift_sim_time", C.c_int, [_VP, _D, _D, _D]),
    ("rift_sim_grid", _VP, [_VP]),
    ("rift_sim_set_source_value", C.c_int, [_VP, C.c_double]),
    ("rift_sim_write_hotstart", C.c_int, [_VP, C.c_char_p]),
    ("rift_sim_read_hotstart", C.c_int, [_VP, C.c_char_p]),
    ("rift_sim_close", C.c_int, [_VP]),
    ("rift_copy_field", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _VP]),
    ("rift_device_alloc", C.c_int, [C.c_int, C.c_size_t, C.POINTER(_VP)]),
    ("rift_device_free", C.c_int, [_VP]),
    ("rift_memcpy", C.c_int, [_VP, _VP, C.c_size_t, C.c_int]),
    ("rift_unet_create", C.c_int, [C.POINTER(UNetDesc), C.POINTER(_VP)]),
    ("rift_unet_destroy", C.c_int, [_VP]),
    ("rift_unet_param_count", C.c_int, [_VP, C.POINTER(C.c_long)]),
    ("rift_unet_get_params", C.c_int, [_VP, _F]),
    ("rift_unet_set_params", C.c_int, [_VP, _F]),
    ("rift_unet_set_loss_weights", C.c_int, [_VP, _F]),
    ("rift_unet_set_adam", C.c_int, [_VP, C.c_float, C.c_float, C.c_float]),
    ("rift_unet_forward", C.c_int, [_VP, _VP, C.c_int, _VP]),
    ("rift_unet_gradient", C.c_int, [_VP, _VP, _VP, C.c_int, _D, _F]),
    ("rift_unet_train_step", C.c_int, [_VP, _VP, _VP, C.c_int, C.c_float, _D]),
    ("rift_unet_flops", C.c_int, [_VP, C.c_int, _D, _D]),
]

_lib = None


class RiftError(RuntimeError):
    pass


def load():
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RiftError(f"{LIB_PATH} is missing: run `make -C rift_amd` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise RiftError(f"{what} failed ({rc}): {load().rift_last_error().decode()}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def kappa(dx, dy):
    return load().rift_kappa(dx, dy)


def geometry(dem, h0=None, n=None, n_const=0.0, K=None):
    """InitSimulation's internal fields from corner rasters (row 0 = south).
    Returns dict with BX (ny, nx+1), BY (ny+1, nx), h0, n, K (ny, nx)."""
    dem = np.ascontiguousarray(dem, dtype=np.float64)
    b_ny, b_nx = dem.shape
    nx, ny = b_nx + 3, b_ny + 3
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (h0, n, K)]
    BX = np.empty((ny, nx + 1))
    BY = np.empty((ny + 1, nx))
    H0 = np.empty((ny, nx)) if h0 is not None else None
    N = np.empty((ny, nx))
    KK = np.empty((ny, nx)) if K is not None else None
    _check(load().rift_geometry(b_nx, b_ny, _dp(dem), _dp(arrs[0]), _dp(arrs[1]), float(n_const),
                                _dp(arrs[2]), _dp(BX), _dp(BY), _dp(H0), _dp(N), _dp(KK)),
           "rift_geometry")
    return {"BX": BX, "BY": BY, "h0": H0, "n": N, "K": KK, "nx": nx, "ny": ny}


def corners_from_faces(BX, BY):
    """Corner elevations (ny-3, nx-3) behind face beds BX (ny, nx+1) and BY
    (ny+1, nx); None when the faces are not means of one corner field."""
    BX = np.ascontiguousarray(BX, dtype=np.float64)
    BY = np.ascontiguousarray(BY, dtype=np.float64)
    ny, nx1 = BX.shape
    nx = nx1 - 1
    D = np.empty((ny - 3, nx - 3))
    rc = load().rift_corners_from_faces(nx, ny, _dp(BX), _dp(BY), _dp(D))
    return D if rc == 0 else None


class DeviceArray:
    """float32 device memory allocated through the library
    (rift_device_alloc); `ptr(i)` is the address of flat element i."""

    def __init__(self, shape, device=0):
        self.shape = tuple(int(v) for v in np.atleast_1d(shape))
        self.size = int(np.prod(self.shape))
        self.device = device
        p = _VP()
        _check(load().rift_device_alloc(device, 4 * max(self.size, 1), C.byref(p)), "rift_device_alloc")
        self._p = p.value

    def ptr(self, i=0):
        if not 0 <= i <= self.size:
            raise RiftError("DeviceArray.ptr: index out of range")
        return self._p + 4 * int(i)

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if offset < 0 or offset + a.size > self.size:
            raise RiftError("DeviceArray.upload: out of range")
        _check(load().rift_memcpy(C.c_void_p(self.ptr(offset)), a.ctypes.data_as(_VP), 4 * a.size, 0),
               "rift_memcpy")

    def download(self, count=None, offset=0, shape=None):
        count = self.size - offset if count is None else int(count)
        if offset < 0 or count < 0 or offset + count > self.size:
            raise RiftError("DeviceArray.download: out of range")
        out = np.empty(count, np.float32)
        _check(load().rift_memcpy(out.ctypes.data_as(_VP), C.c_void_p(self.ptr(offset)), 4 * count, 1),
               "rift_memcpy")
        return out.reshape(shape if shape is not None else ((count,) if count != self.size else self.shape))

    def copy_from(self, other, count, dst_offset=0, src_offset=0):
        if dst_offset + count > self.size or src_offset + count > other.size:
            raise RiftError("DeviceArray.copy_from: out of range")
        _check(load().rift_memcpy(C.c_void_p(self.ptr(dst_offset)), C.c_void_p(other.ptr(src_offset)),
                                  4 * int(count), 2), "rift_memcpy")

    def free(self):
        if getattr(self, "_p", None):
            load().rift_device_free(C.c_void_p(self._p))
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Grid:
    """One simulation domain on one GPU (a rift_grid handle)."""

    def __init__(self, nx, ny, dx, *, dy=None, kappa_=None, courant=0.25, psi=0.0, dtheta=0.0,
                 nodata=-9999.0, flags=TILE_ACCEL, device=0):
        L = load()
        dy = dx if dy is None else dy
        k = L.rift_kappa(dx, dy) if kappa_ is None else kappa_
        self.desc = RiftDesc(nx, ny, dx, dy, k, courant, psi, dtheta, nodata, flags, device)
        h = _VP()
        _check(L.rift_create(C.byref(self.desc), C.byref(h)), "rift_create")
        self.h = h
        self.nx, self.ny, self.flags = nx, ny, flags
        a, b = C.c_int(), C.c_int()
        _check(L.rift_get_tile_dims(self.h, C.byref(a), C.byref(b)), "rift_get_tile_dims")
        self.nbx, self.nby = a.value, b.value
        self.ntiles = self.nbx * self.nby

    @classmethod
    def borrow(cls, handle):
        """A Grid view of a handle owned elsewhere (a Simulation's grid);
        closing the view does not destroy the handle."""
        L = load()
        self = cls.__new__(cls)
        self.desc = RiftDesc()
        _check(L.rift_get_desc(handle, C.byref(self.desc)), "rift_get_desc")
        self.h = handle
        self._owned = False
        self.nx, self.ny, self.flags = self.desc.nx, self.desc.ny, self.desc.flags
        a, b = C.c_int(), C.c_int()
        _check(L.rift_get_tile_dims(self.h, C.byref(a), C.byref(b)), "rift_get_tile_dims")
        self.nbx, self.nby = a.value, b.value
        self.ntiles = self.nbx * self.nby
        return self

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                load().rift_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- setup ---------------------------------------------------------------
    def upload_static(self, BX, BY, n, K=None, h0=None):
        keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (BX, BY, n, K, h0)]
        _check(load().rift_upload_static(self.h, *[_dp(a) for a in keep]), "rift_upload_static")

    def init(self, t0=0.0):
        _check(load().rift_init(self.h, t0), "rift_init")

    def set_series(self, which, time_s, values):
        t = np.ascontiguousarray(time_s, dtype=np.float32)
        v = np.ascontiguousarray(values, dtype=np.float32)
        _check(load().rift_set_series(self.h, which, t.ctypes.data_as(_F), v.ctypes.data_as(_F),
                                      len(t)), "rift_set_series")

    def set_source_cells(self, idx):
        a = np.ascontiguousarray(idx, dtype=np.int32)
        _check(load().rift_set_source_cells(self.h, a.ctypes.data_as(_I), len(a)),
               "rift_set_source_cells")

    def set_source_value(self, q):
        _check(load().rift_set_source_value(self.h, q), "rift_set_source_value")

    def set_map(self, which, plane, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(load().rift_set_map(self.h, which, plane, _dp(a)), "rift_set_map")

    def interp_map(self, which, factor):
        _check(load().rift_interp_map(self.h, which, factor), "rift_interp_map")

    def copy_map(self, which, dst, src):
        _check(load().rift_copy_map(self.h, which, dst, src), "rift_copy_map")

    def map_series(self, which, nslots, deltat, tmax):
        """Device-resident interpolated map series (rift_map_series_create)."""
        _check(load().rift_map_series_create(self.h, which, int(nslots), int(deltat), float(tmax)),
               "rift_map_series_create")

    def load_map(self, which, index, cells):
        """Map `index` (internal grid, (ny, nx) doubles) into its slot."""
        a = np.ascontiguousarray(cells, dtype=np.float64)
        assert a.shape == (self.ny, self.nx), a.shape
        _check(load().rift_map_series_load(self.h, which, int(index), _dp(a)), "rift_map_series_load")

    def start_maps(self, which, t):
        """The series' first p_update at time t (after init / set_time)."""
        _check(load().rift_map_series_start(self.h, which, float(t)), "rift_map_series_start")

    def map_state(self, which):
        """(k0, k1, factor, first missing map index or -1) of the window in force."""
        k0, k1, miss = C.c_int(), C.c_int(), C.c_int()
        f = C.c_double()
        _check(load().rift_map_series_state(self.h, which, C.byref(k0), C.byref(k1), C.byref(f),
                                            C.byref(miss)), "rift_map_series_state")
        return k0.value, k1.value, f.value, miss.value

    # --- stepping ------------------------------------------------------------
    def step(self, n=1):
        _check(load().rift_step(self.h, n), "rift_step")

    def phase(self, name):
        _check(getattr(load(), "rift_" + name)(self.h), "rift_" + name)

    def set_path(self, per_phase):
        """True: one kernel per reference phase; False: the fused step."""
        _check(load().rift_set_option(self.h, 1, 1 if per_phase else 0), "rift_set_option")

    def set_graphs(self, on):
        """Replay the fused step from captured HIP graphs (default on)."""
        _check(load().rift_set_option(self.h, 2, 1 if on else 0), "rift_set_option")

    def set_variant(self, v):
        """Launch variant of the fused step's marching kernels (0..NVARIANTS-1)."""
        _check(load().rift_set_option(self.h, OPT_VARIANT, int(v)), "rift_set_option")

    def set_budget(self, on):
        """FP64 volume budget kernels on (resets the budget) or off."""
        _check(load().rift_set_option(self.h, OPT_MASS_BUDGET, 1 if on else 0), "rift_set_option")

    def set_dt_kernel(self, on):
        """True: the fused step forms dt in a kernel of its own between the
        two marches; False (default): the stage-1 march's last workgroup
        forms it.  Same results bit for bit."""
        _check(load().rift_set_option(self.h, OPT_DT_KERNEL, 1 if on else 0), "rift_set_option")

    def autotune(self, nsteps=16):
        """Times every launch variant on a copy of the current state, selects
        the fastest; returns (best, ms_per_step of each variant)."""
        best = C.c_int()
        ms = np.zeros(NVARIANTS)
        _check(load().rift_autotune(self.h, int(nsteps), C.byref(best), _dp(ms)), "rift_autotune")
        return best.value, ms

    def sync(self):
        _check(load().rift_sync(self.h), "rift_sync")

    # --- readback ------------------------------------------------------------
    @property
    def time(self):
        t, dt = C.c_double(), C.c_double()
        _check(load().rift_get_time(self.h, C.byref(t), C.byref(dt)), "rift_get_time")
        return t.value, dt.value

    def set_time(self, t, dt):
        _check(load().rift_set_time(self.h, t, dt), "rift_set_time")

    def field(self, name):
        out = np.empty((self.ny, self.nx))
        _check(load().rift_download(self.h, FIELDS.index(name), _dp(out)), "rift_download " + name)
        return out

    def copy_field(self, name, dst, i0, j0, w, h):
        """Window [i0, i0+w) x [j0, j0+h) of a stored field into device memory
        at pointer `dst` (h x w float32), ordered on the grid's stream."""
        _check(load().rift_copy_field(self.h, FIELDS.index(name), i0, j0, w, h, C.c_void_p(int(dst))),
               "rift_copy_field " + name)

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(load().rift_upload(self.h, FIELDS.index(name), _dp(a)), "rift_upload " + name)

    def tiles(self):
        wet = np.empty(self.ntiles, dtype=np.uint8)
        act = np.empty(self.ntiles, dtype=np.uint8)
        n = C.c_int()
        _check(load().rift_get_tiles(self.h, wet.ctypes.data_as(_U8), act.ctypes.data_as(_U8),
                                     C.byref(n)), "rift_get_tiles")
        return wet, act, n.value

    def set_tiles(self, wet, act):
        w = np.ascontiguousarray(wet, dtype=np.uint8)
        a = np.ascontiguousarray(act, dtype=np.uint8)
        _check(load().rift_set_tiles(self.h, w.ctypes.data_as(_U8), a.ctypes.data_as(_U8)),
               "rift_set_tiles")

    def active_set(self):
        wet, act, n = self.tiles()
        return set(int(t) for t in np.nonzero(act)[0])

    def tile_speeds(self):
        mx = np.empty(self.ntiles, dtype=np.float32)
        my = np.empty(self.ntiles, dtype=np.float32)
        _check(load().rift_get_tile_speeds(self.h, mx.ctypes.data_as(_F), my.ctypes.data_as(_F)),
               "rift_get_tile_speeds")
        return mx, my

    def rates(self):
        out = np.empty(3, dtype=np.float32)
        _check(load().rift_get_rates(self.h, out.ctypes.data_as(_F)), "rift_get_rates")
        return out

    def sum_interior(self, name):
        out = C.c_double()
        _check(load().rift_sum_interior(self.h, FIELDS.index(name), C.byref(out)), "rift_sum_interior")
        return out.value

    def sum_interior_noalloc(self, name):
        """rift_sum_interior through the library's static scratch (no allocation)."""
        out = C.c_double()
        _check(load().rift_sum_interior_noalloc(self.h, FIELDS.index(name), C.byref(out)),
               "rift_sum_interior_noalloc")
        return out.value

    def budget(self):
        """[stored, infiltrated, total at reset, rain, source, drain, outflow,
        exchange, residual] in m^3 (rift_budget)."""
        out = np.empty(9)
        _check(load().rift_budget(self.h, _dp(out)), "rift_budget")
        return out

    def budget_reset(self):
        _check(load().rift_budget_reset(self.h), "rift_budget_reset")

    def mass_balance(self):
        out = np.empty(6)
        _check(load().rift_mass_balance(self.h, _dp(out)), "rift_mass_balance")
        return out

    def stream(self):
        s = _VP()
        _check(load().rift_get_stream(self.h, C.byref(s)), "rift_get_stream")
        return s.value

    def profile(self, on=True):
        _check(load().rift_profile_enable(self.h, 1 if on else 0), "rift_profile_enable")

    def kernel_stats(self, cls):
        n, ms = C.c_long(), C.c_double()
        _check(load().rift_kernel_stats(self.h, KERNEL_CLASSES.index(cls), C.byref(n), C.byref(ms)),
               "rift_kernel_stats")
        return n.value, ms.value

    def wet_cells(self):
        """(wet-cell updates of the last step, cumulative since init)."""
        a, b = C.c_long(), C.c_long()
        _check(load().rift_wet_cell_updates(self.h, C.byref(a), C.byref(b)), "rift_wet_cell_updates")
        return a.value, b.value

    def interior(self, a):
        return a[2:self.ny - 2, 2:self.nx - 2]


def grid_from_rasters(dem, dx, *, h0=None, n=None, n_const=0.0, K=None, psi=0.0, dtheta=0.0,
                      courant=0.25, flags=TILE_ACCEL, t0=0.0, device=0, per_phase=False):
    """Build, upload and initialise a Grid from corner rasters (row 0 = south)."""
    if h0 is not None:
        flags |= H_INIT
    if K is not None:
        flags |= INFILTRATION
    geo = geometry(dem, h0=h0, n=n, n_const=n_const, K=K)
    g = Grid(geo["nx"], geo["ny"], dx, courant=courant, psi=psi, dtheta=dtheta, flags=flags,
             device=device)
    g.upload_static(geo["BX"], geo["BY"], geo["n"], geo["K"], geo["h0"])
    g.init(t0)
    if per_phase:
        g.set_path(True)
    return g


def map_cells(corners, scale=1.0, allow_nodata=True, nodata=-9999.0):
    """GridSeries::p_interp: the internal-grid map ((b_ny+3, b_nx+3) doubles)
    of a corner map (b_ny, b_nx) -- four-corner means where more than two
    corners hold data, times `scale`; nodata (or 0) elsewhere."""
    c = np.ascontiguousarray(corners, dtype=np.float64)
    b_ny, b_nx = c.shape
    out = np.empty((b_ny + 3, b_nx + 3))
    _check(load().rift_map_cells(b_nx, b_ny, _dp(c), float(scale), int(bool(allow_nodata)), float(nodata),
                                 _dp(out)), "rift_map_cells")
    return out


def read_series(path):
    """CSV series (Source::ReadSource): (time_s, rate) float32 arrays."""
    n = C.c_int()
    _check(load().rift_read_series(os.fsencode(path), None, None, 0, C.byref(n)), "rift_read_series")
    t = np.empty(n.value, np.float32)
    r = np.empty(n.value, np.float32)
    _check(load().rift_read_series(os.fsencode(path), t.ctypes.data_as(_F), r.ctypes.data_as(_F),
                                   n.value, C.byref(n)), "rift_read_series")
    return t, r


def read_raster(path, allow_nodata=False):
    """ESRI-ASCII raster: (data (nrows, ncols) row 0 = south, header dict)."""
    nc, nr = C.c_int(), C.c_int()
    hdr = np.empty(4)
    _check(load().rift_read_raster(os.fsencode(path), int(allow_nodata), C.byref(nc), C.byref(nr),
                                   _dp(hdr), None, 0), "rift_read_raster")
    data = np.empty((nr.value, nc.value))
    _check(load().rift_read_raster(os.fsencode(path), int(allow_nodata), C.byref(nc), C.byref(nr),
                                   _dp(hdr), _dp(data), data.size), "rift_read_raster")
    return data, {"xllcorner": hdr[0], "yllcorner": hdr[1], "cellsize": hdr[2], "nodata": hdr[3]}


def write_grid(path, field, h_xll, h_yll, cellsize):
    """writeGrid: the interior of an internal-grid field (ny, nx)."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    ny, nx = f.shape
    _check(load().rift_write_grid(os.fsencode(path), _dp(f), nx, ny, float(h_xll), float(h_yll),
                                  float(cellsize)), "rift_write_grid")


def write_raster(path, data, xll=0.0, yll=0.0, cellsize=1.0, nodata=-9999.0, fmt="%.10g"):
    """Writes an ESRI-ASCII raster (row 0 of `data` = south) for inputs."""
    data = np.asarray(data, dtype=np.float64)
    nr, nc = data.shape
    with open(path, "w") as f:
        f.write(f"ncols {nc}\nnrows {nr}\nxllcorner {xll!r}\nyllcorner {yll!r}\n"
                f"cellsize {cellsize!r}\nNODATA_value {nodata!r}\n")
        for j in range(nr - 1, -1, -1):
            f.write(" ".join(fmt % v for v in data[j]) + "\n")


class Ensemble:
    """Many domains stepped together (rift_ensemble_*).  `members` are
    dicts of grid_from_rasters arguments (dem, dx, n_const/n, K, psi, dtheta,
    flags, h0, t0); each member is reachable as a borrowed Grid for set-up
    and field access."""

    def __init__(self, members):
        L = load()
        descs, geos, self._kw = [], [], []
        for m in members:
            m = dict(m)
            flags = m.get("flags", TILE_ACCEL)
            if m.get("h0") is not None:
                flags |= H_INIT
            if m.get("K") is not None:
                flags |= INFILTRATION
            geo = geometry(m["dem"], h0=m.get("h0"), n=m.get("n"), n_const=m.get("n_const", 0.0),
                           K=m.get("K"))
            dx = m["dx"]
            descs.append(RiftDesc(geo["nx"], geo["ny"], dx, dx, L.rift_kappa(dx, dx),
                                  m.get("courant", 0.25), m.get("psi", 0.0), m.get("dtheta", 0.0),
                                  -9999.0, flags, m.get("device", 0)))
            geos.append(geo)
            self._kw.append(m)
        arr = (RiftDesc * len(descs))(*descs)
        h = _VP()
        _check(L.rift_ensemble_create(len(descs), arr, C.byref(h)), "rift_ensemble_create")
        self.h = h
        self.members = []
        for i, geo in enumerate(geos):
            g = Grid.borrow(L.rift_ensemble_member(self.h, i))
            g.upload_static(geo["BX"], geo["BY"], geo["n"], geo["K"], geo["h0"])
            g.init(self._kw[i].get("t0", 0.0))
            self.members.append(g)

    def set_tf(self, i, tf):
        _check(load().rift_ensemble_set_tf(self.h, i, float(tf)), "rift_ensemble_set_tf")

    def step(self, n=1):
        _check(load().rift_ensemble_step(self.h, int(n)), "rift_ensemble_step")

    def finished(self):
        v = C.c_int()
        _check(load().rift_ensemble_finished(self.h, C.byref(v)), "rift_ensemble_finished")
        return v.value

    def run(self, check_every=16, max_steps=-1):
        n = C.c_long()
        _check(load().rift_ensemble_run(self.h, int(check_every), int(max_steps), C.byref(n)),
               "rift_ensemble_run")
        return n.value

    def close(self):
        if getattr(self, "h", None):
            for g in self.members:
                g.close()
            h, self.h = self.h, None
            load().rift_ensemble_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Config:
    """The keyword file (ConfigFile, config.h:48-209) through the C ABI."""

    def __init__(self, path):
        h = _VP()
        _check(load().rift_config_open(os.fsencode(path), C.byref(h)), "rift_config_open")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().rift_config_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def has(self, key):
        return bool(load().rift_config_has(self.h, key.encode()))

    def is_number(self, key):
        v = C.c_int()
        _check(load().rift_config_is_number(self.h, key.encode(), C.byref(v)), "rift_config_is_number")
        return bool(v.value)

    def string(self, key):
        n = C.c_int()
        _check(load().rift_config_string(self.h, key.encode(), None, 0, C.byref(n)), "rift_config_string")
        buf = C.create_string_buffer(n.value + 1)
        _check(load().rift_config_string(self.h, key.encode(), buf, n.value + 1, None), "rift_config_string")
        return buf.value.decode()

    def double(self, key):
        v = C.c_double()
        _check(load().rift_config_double(self.h, key.encode(), C.byref(v)), "rift_config_double")
        return v.value

    def bool(self, key):
        v = C.c_int()
        _check(load().rift_config_bool(self.h, key.encode(), C.byref(v)), "rift_config_bool")
        return bool(v.value)

    def complex(self, key):
        a, b = C.c_double(), C.c_double()
        _check(load().rift_config_complex(self.h, key.encode(), C.byref(a), C.byref(b)), "rift_config_complex")
        return complex(a.value, b.value)


class Simulation:
    """The Simulator driver (simulator.h:14-173) through the C ABI:
    open = OpenSimulation, step = RunSimulation, close = CloseSimulation."""

    def __init__(self, config_path, flow_rate=0.0):
        h = _VP()
        _check(load().rift_sim_open(os.fsencode(config_path), float(flow_rate), C.byref(h)), "rift_sim_open")
        self.h = h
        self.grid = Grid.borrow(load().rift_sim_grid(self.h))

    def step(self):
        t = C.c_double()
        _check(load().rift_sim_step(self.h, C.byref(t)), "rift_sim_step")
        return t.value

    def run(self, t_end=-1.0):
        _check(load().rift_sim_run(self.h, float(t_end)), "rift_sim_run")

    def time(self):
        t, dt, tf = C.c_double(), C.c_double(), C.c_double()
        _check(load().rift_sim_time(self.h, C.byref(t), C.byref(dt), C.byref(tf)), "rift_sim_time")
        return t.value, dt.value, tf.value

    def set_source_value(self, q):
        _check(load().rift_sim_set_source_value(self.h, float(q)), "rift_sim_set_source_value")

    def write_hotstart(self, path):
        _check(load().rift_sim_write_hotstart(self.h, os.fsencode(path)), "rift_sim_write_hotstart")

    def read_hotstart(self, path):
        _check(load().rift_sim_read_hotstart(self.h, os.fsencode(path)), "rift_sim_read_hotstart")

    def close(self):
        if getattr(self, "h", None):
            self.grid.close()
            h, self.h = self.h, None
            _check(load().rift_sim_close(h), "rift_sim_close")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
