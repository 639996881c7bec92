"""rift_sum_interior_noalloc (static device scratch, no allocation) against
rift_sum_interior through the C ABI: two calls in a row on each field, the
second on a smaller domain (stale partials of the first call must not enter
it), agreeing to 1e-12 relative (FP64 sums of the same FP32 values in
another order); a domain beyond the scratch's 8192^2 interior cells is
refused with RIFT_E_UNSUPPORTED before any device work."""
import numpy as np
import pytest

from rift_amd import synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_noalloc_sum_matches_the_allocating_sum(rift):
    cfg = synth.rain_basin(1024, seed=5)
    big = rift.grid_from_rasters(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"],
                                 psi=cfg["psi"], dtheta=cfg["dtheta"],
                                 flags=rift.TILE_ACCEL | rift.RAIN_AVG)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    big.set_series(rift.SER_HYETO, t, v)
    big.step(400)
    small_cfg = synth.dam_break(96)
    small = rift.grid_from_rasters(small_cfg["dem"], small_cfg["dx"], h0=small_cfg["h0"])
    small.step(30)
    print("no-alloc sum: 1024^2 rain basin then 96^2 dam break", flush=True)
    for g in (big, small, big):
        for f in ("depth", "hu", "hv"):
            a = g.sum_interior(f)
            b = g.sum_interior_noalloc(f)
            assert _rel(a, b) <= 1e-12, (g.nx, f, a, b)
            # the same call again gives the same bits (fixed summation order)
            assert g.sum_interior_noalloc(f) == b
        h = g.interior(g.field("depth"))
        assert _rel(g.sum_interior_noalloc("depth"), float(h.sum(dtype=np.float64))) <= 1e-12
    big.close()
    small.close()


def test_noalloc_sum_refuses_a_domain_beyond_the_scratch(rift):
    g = rift.Grid(8200 + 4, 8200 + 4, 5.0)   # 8200^2 interior cells > 8192^2
    with pytest.raises(rift.RiftError, match="static scratch"):
        g.sum_interior_noalloc("depth")
    g.close()
