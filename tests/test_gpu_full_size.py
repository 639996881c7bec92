"""The full benchmark size: the 2048 x 2048 rain basin (rain, Green-Ampt,
Manning, tile acceleration; rift_amd.synth.rain_basin(2048, seed=2), the
workload bench.py times) stepped on the GPU until every tile is active.

Parity: the FP64 oracle continues from the GPU's state as a hotstart does
(the interior depth, hu, hv and F, tile flags, t and dt copied over, then
the outflow boundary applied, simulator.cu:725-731) and both advance side by
side; every step the dt agrees (relative 5e-3, the thin-sheet-flow bound of
test_gpu_parity) and the wet and active tile sets are identical, and at the
end the depth agrees to 1e-4 m.  The ghost cells the GPU reports for w are
the reference's (w = BC of the adjacent interior cell + critical depth).

Launch variants: the marching kernels' launch shapes (waves per SIMD, the
XCD dealing of the row shares, the least share) change only which wave
computes which row, so two grids stepped from the same state with different
variants stay bit-identical at this size, where every resident wave holds
work."""
import numpy as np
import pytest

import oracle_lib as O
from rift_amd import synth
from test_gpu_parity import _run_compare

pytestmark = pytest.mark.gpu

SIDE = 2048
WARMUP = 160


def _grid(rift, cfg):
    g = rift.grid_from_rasters(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"],
                               psi=cfg["psi"], dtheta=cfg["dtheta"],
                               flags=rift.TILE_ACCEL | rift.RAIN_AVG)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    g.set_series(rift.SER_HYETO, t, v)
    return g


@pytest.fixture(scope="module")
def basin():
    return synth.rain_basin(SIDE, seed=2)


def test_full_size_against_the_oracle(rift, basin):
    cfg = basin
    print("rain basin", SIDE, "warm-up", WARMUP, flush=True)
    g = _grid(rift, cfg)
    g.step(WARMUP)
    wet, act, n_act = g.tiles()
    assert n_act == g.ntiles          # every tile active: the dense marching path
    o = O.Oracle(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"], psi=cfg["psi"],
                 dtheta=cfg["dtheta"], flags=O.TILE_ACCEL | O.RAIN_AVG)
    o.set_series_hours(O.SER_HYETO, cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    o.set_field("w", g.field("depth") + o.field("bc"))
    for f in ("hu", "hv", "F"):
        o.set_field(f, g.field(f))
    o.phase("apply_boundaries")
    # the GPU's w, ghost cells included, is the reference's up to the FP32
    # rounding of the beds (ulp of 60 m = 3.8e-6 m); the 2 x 2 corner ghosts
    # are read by neither pass and written by neither boundary
    used = np.ones((g.ny, g.nx), bool)
    for js in (slice(0, 2), slice(-2, None)):
        for is_ in (slice(0, 2), slice(-2, None)):
            used[js, is_] = False
    np.testing.assert_allclose(g.field("w")[used], o.field("w")[used], rtol=0, atol=2e-5)
    o.set_tiles(wet, np.where(act > 0, np.arange(g.ntiles), -np.arange(g.ntiles)))
    t, dt = g.time
    o.set_time(t, dt)
    steps, dt_err = _run_compare(o, g, nsteps=8, dt_rtol=5e-3)
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    du = max(np.abs(g.interior(g.field(f)) - o.interior(o.field(f))).max() for f in ("hu", "hv"))
    print(f"steps {steps} from t = {t:.1f} s: max|dh| {dh:.2e} m, max|dq| {du:.2e} m^2/s, "
          f"dt rel {dt_err:.1e}", flush=True)
    assert dh <= 1e-4, dh
    assert g.interior(g.field("depth")).max() > 0.01   # 4.9 cm at this time
    g.close()


def test_full_size_launch_variants_bit_identical(rift, basin):
    g0, g1 = _grid(rift, basin), _grid(rift, basin)
    g0.step(WARMUP)
    g1.step(WARMUP)
    g0.set_variant(0)
    g1.set_variant(15)
    g0.step(24)
    g1.step(24)
    assert g0.time == g1.time
    for f in ("depth", "hu", "hv", "F"):
        np.testing.assert_array_equal(g0.field(f), g1.field(f), err_msg=f)
    np.testing.assert_array_equal(g0.tiles()[0], g1.tiles()[0])
    g0.close()
    g1.close()
