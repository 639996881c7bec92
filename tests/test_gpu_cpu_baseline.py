"""The CPU baseline (the FP32 OpenMP transcription of the per-cell update,
oracle/kp_oracle.c built with ORC_FLOAT) against the HIP path: both FP32,
the same scheme and tile logic.  They differ only in how the FP32
operations are ordered and rounded (explicit FMAs, the hardware reciprocal,
square root, log and exp on the device; libm on the host).  On the thin
sheet flow of a rain basin with Green-Ampt infiltration and Manning friction
the tile sets match every step and dt agrees to 5e-3 relative (the bound of
test_gpu_parity for thin sheet flow: the face depth h + (BC - B_face) +
slope is a small difference of ~0.5 m terms that FP32 resolves to ~1e-7 m,
and the two time series drift apart by it); the depth after 60 steps agrees
to the same 1e-4 m the device keeps to the FP64 oracle (measured 4.8e-5 m)."""
import numpy as np
import pytest

import oracle_lib as O
from rift_amd import synth

pytestmark = pytest.mark.gpu


def test_fp32_cpu_baseline_tracks_the_device(rift):
    cfg = synth.rain_basin(128, seed=7)
    print("rain basin 128, fp32 CPU baseline vs device", flush=True)
    o = O.Oracle(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"], psi=cfg["psi"],
                 dtheta=cfg["dtheta"], flags=O.TILE_ACCEL | O.RAIN_AVG, fp32=True)
    o.set_series_hours(O.SER_HYETO, cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    g = rift.grid_from_rasters(cfg["dem"], cfg["dx"], n_const=cfg["n_const"], K=cfg["K"],
                               psi=cfg["psi"], dtheta=cfg["dtheta"],
                               flags=rift.TILE_ACCEL | rift.RAIN_AVG)
    t, v = synth.series_hours_to_seconds(cfg["hyeto_hours"], cfg["hyeto_mmhr"])
    g.set_series(rift.SER_HYETO, t, v)
    worst_dt = 0.0
    for k in range(60):
        o.step()
        g.step()
        (to, dto), (tg, dtg) = o.time, g.time
        worst_dt = max(worst_dt, abs(dtg - dto) / dto)
        assert abs(dtg - dto) <= 5e-3 * dto, (k, dto, dtg)
        np.testing.assert_array_equal(g.tiles()[0], o.tiles()[0])
    dh = np.abs(g.interior(g.field("depth")) - o.interior(o.depth())).max()
    print(f"max|dh| {dh:.2e} m, dt rel {worst_dt:.1e}", flush=True)
    assert dh <= 1e-4, dh
    assert g.interior(g.field("depth")).max() > 0.0
    g.close()
