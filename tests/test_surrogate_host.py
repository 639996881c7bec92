"""Host side of the surrogate workflow: scenario inputs and features."""
import numpy as np

from rift_amd import surrogate as S


def test_scenarios_are_seeded_and_in_range():
    a, b = S.scenario(5001, side=64), S.scenario(5001, side=64)
    np.testing.assert_array_equal(a["dem"], b["dem"])
    np.testing.assert_array_equal(a["K"], b["K"])
    assert a["n_const"] == b["n_const"] and 0.02 <= a["n_const"] <= 0.08
    c = S.scenario(5002, side=64)
    assert not np.array_equal(a["dem"], c["dem"])
    assert a["dem"].shape == (65, 65) and a["K"].shape == (65, 65)
    assert a["mmhr"][0] == 0 and a["mmhr"][2] == 0 and a["mmhr"][1] > 0


def test_features_of_a_plane():
    """A tilted plane: slopes are the plane's gradient (scaled by 1/0.05), the
    relief against local means vanishes away from the edges, elevation runs
    from 0 upwards, and the scalar channels are constant."""
    side = 64
    sc = S.scenario(5003, side=side)
    y, x = np.mgrid[0:side + 1, 0:side + 1].astype(np.float64)
    sc["dem"] = 0.3 * x * sc["dx"] * 0.01 + 0.2 * y * sc["dx"] * 0.01 + 7.0   # 0.3 % and 0.2 %
    f = S.features(sc)
    assert f.shape == (len(S.FEATURES), side, side) and f.dtype == np.float32
    inner = (slice(17, -17), slice(17, -17))
    np.testing.assert_allclose(f[3][1:-1, 1:-1], 0.003 / 0.05, rtol=1e-5)
    np.testing.assert_allclose(f[4][1:-1, 1:-1], 0.002 / 0.05, rtol=1e-5)
    assert np.abs(f[1][inner]).max() < 1e-6 and np.abs(f[2][inner]).max() < 1e-6
    assert f[0].min() == 0.0
    for c in (5, 6):
        assert np.ptp(f[c]) == 0.0
    assert abs(f[6, 0, 0] - np.log(sc["mmhr"].max() / 50.0) / 0.5) < 1e-6


def test_asset_depth_takes_the_central_window():
    m = np.zeros((2, 32, 32), np.float32)
    m[0, 16, 16] = 0.7
    m[1, 0, 0] = 5.0            # outside the window
    m[1, 12, 19] = 0.2          # corner of the window (cells 12..19)
    np.testing.assert_allclose(S.asset_depth(m), [0.7, 0.2])


def test_error_stats():
    t = np.zeros((3, 2, 16, 16), np.float32)
    p = t.copy()
    p[:, 0] += 0.05
    e = S.error_stats(p, t)
    assert abs(e["depth_mae_m"] - 0.05) < 1e-7 and abs(e["depth_rmse_m"] - 0.05) < 1e-7
    assert e["arrival_mae_h"] == 0.0
