"""Edge cases of the HIP path against the oracle: ragged domains whose sides
are no multiple of the 16 x 14 tile or of the 60-column marching strip
(including an 8 x 8 interior and a 250-column, 37-row strip), a dry domain
with nothing to move, and the size limit of rift_create.  Tolerances as in
test_gpu_parity (depth 1e-4 m, dt 5e-3 relative on thin sheet flow, tile
sets identical every step)."""
import numpy as np
import pytest

import oracle_lib as O
from rift_amd import synth
from test_gpu_parity import _pair, _run_compare

pytestmark = pytest.mark.gpu

RAGGED = [(74, 142), (38, 251), (9, 9), (15, 17), (130, 61)]   # DEM points (rows, columns)


def _dem(shape, seed):
    n = max(shape) - 1
    return 30.0 * synth.fractal_dem(n, 0.6, seed)[: shape[0], : shape[1]]


@pytest.mark.parametrize("per_phase", [False, True], ids=["fused", "per_phase"])
@pytest.mark.parametrize("shape", RAGGED, ids=[f"{a}x{b}" for a, b in RAGGED])
def test_ragged_rain_basin(rift, shape, per_phase):
    print("ragged", shape, per_phase, flush=True)
    dem = _dem(shape, 11)
    K = 5.0 * np.exp(0.8 * synth.normal(12, dem.size).reshape(dem.shape))
    o, g = _pair(rift, dem, 10.0, n_const=0.03, K=K, psi=0.11, dtheta=0.3,
                 oflags=O.TILE_ACCEL | O.RAIN_AVG, per_phase=per_phase)
    hours = np.array([0.0, 0.05, 0.2], np.float32)
    mmhr = np.array([0.0, 150.0, 0.0], np.float32)
    o.set_series_hours(O.SER_HYETO, hours, mmhr)
    t, v = synth.series_hours_to_seconds(hours, mmhr)
    g.set_series(rift.SER_HYETO, t, v)
    _run_compare(o, g, nsteps=80, dt_rtol=5e-3)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    assert g.interior(g.field("depth")).max() > 0
    g.close()


@pytest.mark.parametrize("shape", [(57, 200), (200, 31)], ids=["wide", "tall"])
def test_ragged_dam_break(rift, shape):
    ny, nx = shape
    dem = np.zeros(shape)
    h0 = np.where(np.arange(nx)[None, :] < nx // 3, 1.0, 0.05) * np.ones((ny, 1))
    o, g = _pair(rift, dem, 1.0, h0=h0, n_const=0.01)
    _run_compare(o, g, nsteps=120)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    assert dh <= 1e-4, dh
    g.close()


def test_dry_domain(rift):
    """No water and no forcing: the tile sets and dt follow the oracle and the
    depth stays zero.  The momenta do not stay exactly zero in either: on a
    sloping bed a dry cell's face depth BC - B_face + slope can be positive,
    so the reconstruction leaves tiny pressure fluxes (grid.cu:91-116)."""
    dem = _dem((33, 45), 13)
    o, g = _pair(rift, dem, 5.0, n_const=0.03)
    _run_compare(o, g, nsteps=10)
    assert np.abs(g.interior(g.field("depth"))).max() == 0.0
    assert np.abs(o.interior(o.depth())).max() == 0.0
    for f in ("hu", "hv"):
        d = np.abs(g.interior(g.field(f)) - o.interior(o.field(f))).max()
        assert d <= 1e-7, (f, d)
    g.close()


def test_oversized_domain_is_rejected(rift):
    """A field plane must hold fewer than 2^30 cells; larger domains are
    refused before any allocation."""
    with pytest.raises(rift.RiftError, match="too large"):
        rift.Grid(40003, 40003, 1.0)


def test_beds_that_are_no_corner_means_run_the_per_phase_path(rift):
    """Face beds that are not two-point means of one corner field (no
    reference geometry gives them) keep the uploaded planes: rift_step runs
    the per-phase kernels, bit for bit as when that path is asked for, and
    the autotuner (which times the fused step) refuses the grid."""
    dem = _dem((40, 52), 4)
    geo = rift.geometry(dem, n_const=0.03)
    BX = geo["BX"].copy()
    BX[20, 25] += 0.05   # one interior face off its corners
    assert rift.corners_from_faces(BX, geo["BY"]) is None
    h0 = np.maximum(18.0 - dem, 0.0)
    geo_h = rift.geometry(dem, h0=h0, n_const=0.03)
    grids = []
    for per_phase in (False, True):
        g = rift.Grid(geo["nx"], geo["ny"], 5.0, flags=rift.TILE_ACCEL | rift.H_INIT)
        g.upload_static(BX, geo["BY"], geo["n"], None, geo_h["h0"])
        g.init(0.0)
        if per_phase:
            g.set_path(True)
        g.step(60)
        grids.append(g)
    a, b = grids
    assert a.time == b.time
    for f in ("depth", "hu", "hv"):
        np.testing.assert_array_equal(a.field(f), b.field(f))
    assert np.abs(a.field("hu")).max() > 0
    with pytest.raises(rift.RiftError, match="corner"):
        a.autotune(4)
    for g in grids:
        g.close()
