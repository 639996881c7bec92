"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors or tests for the step (its only test
target, GridSeriesTest.cpp, asserts nothing).  Each oracle function is pinned
here by a check chosen so that a plausible slip (dropped term, wrong sign,
transposed operand) fails:
  signf/minmod/limiter   exact bit patterns of the low-word rule
  velocity/kp_flux       closed forms and consistency (F(U,U) = physical flux)
  construct              positivity branches, exact reconstruction of a line
  face fluxes            x/y symmetry, lake at rest, Stoker dam break
  sources                uniform rain rise, semi-implicit friction decay,
                         Green-Ampt quadratic root
  tiles                  brute force over every wet pattern of small grids
  dt                     the ComputeTimestep formula on a known speed
"""
import itertools
import math
import struct

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle_lib as O
from rift_amd import synth

G = float(np.float32(9.80665))
EPS = float(np.float32(1.19209e-07))
THETA = float(np.float32(1.3))


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


# --- signf / minmod / limiter (grid.cu:62-80) -------------------------------

def test_signf_low_word_rule():
    L = O.lib()
    # "short" doubles have a zero low word: sign 0, whatever their sign
    for v in (0.0, -0.0, 0.5, -0.5, 1.0, -1.0, -3.0, 2.0**-20, 1024.0):
        assert L.orc_signf(v) == 0.0, v
    # 0.9 = 0x3FECCCCCCCCCCCCD: low word 0xCCCCCCCD has bit 31 set -> 1 + 2^-21
    assert _bits(0.9) & 0xFFFFFFFF == 0xCCCCCCCD
    assert L.orc_signf(0.9) == 1.0 + 2.0**-21
    assert L.orc_signf(-0.9) == 1.0 + 2.0**-21     # never negative
    # low word without bit 31 -> exactly 1
    v = struct.unpack("<d", struct.pack("<Q", 0x3FF0000000000001))[0]
    assert L.orc_signf(v) == 1.0
    assert L.orc_signf(-v) == 1.0
    # only bits 0..30 are tested: bit 31 alone reads as zero
    v = struct.unpack("<d", struct.pack("<Q", 0x3FF0000080000000))[0]
    assert L.orc_signf(v) == 0.0


def test_minmod_sign_blind_cases():
    L = O.lib()
    a, b, c = 0.3, -0.7, 0.11          # generic doubles, mixed signs
    k = L.orc_signf(a) * (L.orc_signf(a) + L.orc_signf(b)) * (L.orc_signf(b) + L.orc_signf(c)) * 0.25
    expect = k * float(np.float32(min(abs(np.float32(a)), abs(np.float32(b)), abs(np.float32(c)))))
    assert L.orc_minmod(O.FAITHFUL, a, b, c) == expect
    assert L.orc_minmod(O.FAITHFUL, a, b, c) > 0          # sign-blind: positive
    assert L.orc_minmod(O.TEXTBOOK, a, b, c) == 0.0        # textbook: mixed -> 0
    assert L.orc_minmod(O.TEXTBOOK, -0.3, -0.7, -0.11) == -0.11
    assert L.orc_minmod(O.TEXTBOOK, 0.3, 0.7, 0.11) == 0.11
    # short-double middle operand: s(b) = 0 -> factor 1/4 * s(a)^2 s(c)
    r = L.orc_minmod(O.FAITHFUL, 0.3, 0.5, 0.7)
    sa, sc = L.orc_signf(0.3), L.orc_signf(0.7)
    assert r == 0.25 * sa * sa * sc * float(np.float32(0.3))
    # short-double last operand: factor 1/2
    r = L.orc_minmod(O.FAITHFUL, 0.3, 0.7, 0.5)
    sa, sb = L.orc_signf(0.3), L.orc_signf(0.7)
    assert r == 0.25 * sa * (sa + sb) * sb * float(np.float32(0.3))
    # any exact zero -> 0
    assert L.orc_minmod(O.FAITHFUL, 0.3, 0.0, 0.7) == 0.0


def test_limiter_reconstructs_a_line_textbook():
    L = O.lib()
    # on a linear profile the textbook limiter returns half the slope
    # (theta > 1 so the central difference wins)
    for s in (0.37, -1.3, 2.5e-3):
        got = L.orc_limiter(O.TEXTBOOK, 1.0, 1.0 + s, 1.0 - s)
        assert got == pytest.approx(0.5 * s, rel=1e-12)
    # faithful: the same magnitude, always positive
    got = L.orc_limiter(O.FAITHFUL, 1.0, 1.0 - 0.37, 1.0 + 0.37)
    assert got == pytest.approx(0.5 * 0.37, rel=1e-6) and got > 0


# --- velocity and numerical flux --------------------------------------------

def test_velocity_desingularisation():
    L = O.lib()
    kap = L.orc_kappa(1.0, 1.0)
    assert kap == float(np.float32(math.sqrt(np.float32(0.01))))
    # deep water: u = q/h
    for h, q in ((1.0, 0.3), (2.5, -1.2)):
        assert L.orc_velocity(O.TEXTBOOK, h, q, kap) == pytest.approx(q / h, rel=1e-14)
    # shallow: u = 2 h q / (h^2 + kappa)
    h, q = 0.01, 1e-4
    assert L.orc_velocity(O.TEXTBOOK, h, q, kap) == pytest.approx(2 * h * q / (h * h + kap), rel=1e-14)
    assert L.orc_velocity(O.TEXTBOOK, 0.0, 1.0, kap) == 0.0


def test_kappa_per_cell_size():
    L = O.lib()
    for dx, want in ((1.0, 0.10000000149), (5.0, 0.22360679507), (10.0, 0.31622776389)):
        assert L.orc_kappa(dx, dx) == pytest.approx(want, abs=1e-10)


def test_kp_flux_consistency():
    L = O.lib()
    # equal states: H = F regardless of the speeds
    assert L.orc_kp_flux(2.0, 3.0, 2.0, 3.0, -1.5, 2.5) == pytest.approx(3.0, rel=1e-15)
    # upwind limit cm = 0: H = F(left)
    assert L.orc_kp_flux(1.0, 3.0, 2.0, 7.0, 0.0, 2.5) == pytest.approx(3.0, rel=1e-15)
    # cp = 0: H = F(right)
    assert L.orc_kp_flux(1.0, 3.0, 2.0, 7.0, -2.0, 0.0) == pytest.approx(7.0, rel=1e-15)
    # closed form
    Uij, Fij, Ukh, Fkh, cm, cp = 0.4, 1.1, 0.9, 2.3, -1.7, 3.1
    want = (cp * Fij - cm * Fkh) / (cp - cm) + cp * cm / (cp - cm) * (Ukh - Uij)
    assert L.orc_kp_flux(Uij, Fij, Ukh, Fkh, cm, cp) == pytest.approx(want, rel=1e-15)


def _face(mode, dir_, left, right, B, kappa):
    L = O.lib()
    out = np.zeros(4)
    l = np.asarray(left, float)
    r = np.asarray(right, float)
    L.orc_face(mode, dir_, l.ctypes.data_as(O._D), r.ctypes.data_as(O._D), B, kappa,
               out.ctypes.data_as(O._D))
    return out


def test_face_flux_physical_flux_for_equal_states():
    kap = 0.1
    w, hu, hv, B = 1.7, 0.6, -0.2, 0.3
    h = w - B
    out = _face(O.TEXTBOOK, 0, (w, hu, hv), (w, hu, hv), B, kap)
    u = 2 * h * hu / (h * h + max(h * h, kap))
    v = 2 * h * hv / (h * h + max(h * h, kap))
    np.testing.assert_allclose(out[:3], [h * u, h * u * u + 0.5 * G * h * h, h * u * v], rtol=1e-12)
    assert out[3] == pytest.approx(abs(u) + math.sqrt(G * h), rel=1e-12)
    # y-direction: momentum components swap roles
    outy = _face(O.TEXTBOOK, 1, (w, hv, hu), (w, hv, hu), B, kap)
    np.testing.assert_allclose([outy[0], outy[2], outy[1]], out[:3], rtol=1e-12)


def test_face_flux_dry_sides_and_nan_rule():
    kap = 0.1
    # both sides dry: no flux, no speed
    out = _face(O.FAITHFUL, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, kap)
    assert np.all(out == 0.0)
    # a negative-depth side drops out of the speed bounds (sqrtf -> NaN)
    out = _face(O.FAITHFUL, 0, (-0.1, 0.0, 0.0), (1.0, 0.0, 0.0), 0.0, kap)
    assert out[3] == pytest.approx(math.sqrt(G), rel=1e-6)
    assert np.all(np.isfinite(out))


# --- construct (positivity) ---------------------------------------------------

def test_construct_positivity_branches():
    L = O.lib()
    out = np.zeros(6)

    def cons(mode, c3, Bp, Bm):
        a = np.asarray(c3, float)
        L.orc_construct(mode, a.ctypes.data_as(O._D), Bp, Bm, out.ctypes.data_as(O._D))
        return out.copy()

    # lake at rest: w constant, zero slope, faces equal w
    o = cons(O.FAITHFUL, [2.0, 0, 0, 2.0, 0, 0, 2.0, 0, 0], 1.5, 1.0)
    np.testing.assert_array_equal(o[[0, 3]], [2.0, 2.0])
    # east face below its bed: w_E = B_E, w_W = wbar - B_E (mass kept)
    w, Bp, Bm = 1.0, 1.5, 0.0
    o = cons(O.TEXTBOOK, [w, 0, 0, 2.0, 0, 0, 0.5, 0, 0], Bp, Bm)
    wz = 0.5 * min(THETA * (2.0 - w), 0.5 * (2.0 - 0.5), THETA * (w - 0.5))
    assert o[0] == Bp and o[3] == pytest.approx((w + wz) + (w - wz) - Bp, rel=1e-15)
    assert o[0] + o[3] == pytest.approx(2 * w, rel=1e-15)


# --- whole-step invariants ------------------------------------------------------

def _flat(n=48):
    return np.zeros((n + 1, n + 1))


def test_lake_at_rest_fractal_bed():
    # well-balancing: w constant, u = 0 over a rough bed gives dU = 0 (the
    # hydrostatic flux difference cancels the bed-slope source) away from the
    # outflow ghosts, in both limiter modes; a wrong sign of S1/S2 fails
    from rift_amd import synth
    dem = 2.0 * synth.fractal_dem(32, 0.6, 11)
    h0 = 3.0 - dem                         # corner depths of the surface w = 3
    for mode in (O.FAITHFUL, O.TEXTBOOK):
        o = O.Oracle(dem, 1.0, h0=h0, mode=mode)
        w = o.interior(o.field("w"))
        assert np.abs(w - 3.0).max() < 1e-14
        o.phase("grow")
        o.phase("compute_fluxes")
        c = slice(3, -3)
        for f in ("dw", "dhu", "dhv"):
            d = o.interior(o.field(f))[c, c]
            assert np.abs(d).max() < 1e-11, (mode, f, np.abs(d).max())
        # the bed-slope source alone is large: it is the flux that cancels it
        S1 = -G * o.interior(o.depth()) * np.diff(o.faces()[0], axis=1)[2:-2, 2:-2]
        assert np.abs(S1[c, c]).max() > 1.0


def test_uniform_rain_rise_closed_form():
    # flat bed, uniform rain r = 36 mm/hr = 1e-5 m/s: Heun of a constant
    # source adds r*dt per step; the first step's r*1e-6 m is below epsilon and
    # is clamped away, so h = r (t - 1e-6) away from the outflow boundaries
    dem = _flat(128)
    o = O.Oracle(dem, 1.0, flags=O.TILE_ACCEL | O.RAIN_AVG, mode=O.FAITHFUL)
    o.set_series_hours(O.SER_HYETO, [0.0, 10.0], [36.0, 36.0])
    for _ in range(12):
        o.step()
    t, dt = o.time
    r = o.rates()[0]
    assert r == np.float32(np.float32(36.0) / np.float32(3.6e6))
    h = o.interior(o.depth())
    assert h[64, 64] == pytest.approx(float(r) * (t - float(np.float32(1e-6))), rel=1e-12)


def test_friction_decay_semi_implicit():
    # flat tank, uniform depth and velocity, Manning n: stage 1 divides hu by
    # (1 - dt G); compare one Euler step with the closed form
    n = 0.05
    dem = _flat(32)
    h0 = np.full_like(dem, 1.0)
    o = O.Oracle(dem, 1.0, h0=h0, n_const=n, flags=O.TILE_ACCEL | O.EULER, mode=O.TEXTBOOK)
    hu0 = np.zeros((o.ny, o.nx))
    hu0[:, :] = 0.5
    o.set_field("hu", hu0)
    o.phase("grow")
    o.phase("compute_fluxes")
    o.phase("compute_timestep")
    t, dt = o.time
    Gc = o.interior(o.field("G"))[16, 16]
    kap = O.lib().orc_kappa(1.0, 1.0)
    want_G = -4 * G * n * n * 0.5 * 1.0 / (1.0 + max(1.0, kap)) ** 2
    assert Gc == pytest.approx(want_G, rel=1e-6)
    o.phase("integrate_1")
    hu = o.interior(o.field("hu"))[16, 16]
    assert hu == pytest.approx(0.5 / (1.0 - dt * want_G), rel=1e-6)


def test_green_ampt_root():
    # K, F, psi*dtheta given: the infiltration rate is the positive root of
    # dt f^2 - (K dt - 2F) f - 2 K (F + psi dtheta) = 0
    dem = _flat(32)
    h0 = np.full_like(dem, 0.5)
    Kmm = np.full_like(dem, 18.0)     # mm/hr
    o = O.Oracle(dem, 1.0, h0=h0, K=Kmm, psi=0.11, dtheta=0.3, mode=O.TEXTBOOK)
    F = np.full((o.ny, o.nx), 0.01)
    o.set_field("F", F)
    o.set_time(5.0, 0.2)
    o.phase("grow")
    o.phase("compute_fluxes")
    f = o.interior(o.field("dF"))[16, 16]
    K = 18.0 / 3.6e6
    dt, Fv, pd = 0.2, 0.01, 0.11 * 0.3
    root = brentq(lambda x: dt * x * x - (K * dt - 2 * Fv) * x - 2 * K * (Fv + pd), 0.0, 1.0)
    assert f == pytest.approx(root, rel=1e-9)


def test_critical_depth_boundary():
    dem = _flat(32)
    h0 = np.full_like(dem, 1.0)
    o = O.Oracle(dem, 1.0, h0=h0, mode=O.TEXTBOOK)
    hu = np.zeros((o.ny, o.nx))
    hu[:, 2] = 0.8
    o.set_field("hu", hu)
    o.phase("apply_boundaries")
    w = o.field("w")
    Hc = (0.8 * 0.8 / G) ** float(np.float32(1.0) / np.float32(3.0))   # 1.f/3.f, grid.cu:1342
    assert w[10, 0] == pytest.approx(Hc, rel=1e-12) and w[10, 1] == w[10, 0]
    assert o.field("hu")[10, 0] == 0.8


def test_xy_symmetry_of_the_step():
    # a transposed initial condition gives the transposed solution (except
    # the y boundary's missing epsilon clamp, not exercised by a wet domain)
    from rift_amd import synth
    dem = 1.5 * synth.fractal_dem(32, 0.6, 5)
    h0 = 1.0 + 0.3 * synth.fractal_dem(32, 0.5, 6)
    a = O.Oracle(dem, 1.0, h0=h0, n_const=0.03, mode=O.TEXTBOOK)
    b = O.Oracle(dem.T.copy(), 1.0, h0=h0.T.copy(), n_const=0.03, mode=O.TEXTBOOK)
    a.step(15)
    b.step(15)
    assert a.time == pytest.approx(b.time, rel=1e-13)
    np.testing.assert_allclose(a.depth(), b.depth().T, rtol=0, atol=1e-12)
    np.testing.assert_allclose(a.field("hu"), b.field("hv").T, rtol=0, atol=1e-12)


def _stoker(x, t, x0, hl, hr):
    cl = math.sqrt(G * hl)
    f = lambda hm: 2 * (cl - math.sqrt(G * hm)) - (hm - hr) * math.sqrt(G / 2 * (hm + hr) / (hm * hr))
    hm = brentq(f, hr, hl)
    um = 2 * (cl - math.sqrt(G * hm))
    cm = math.sqrt(G * hm)
    s = hm * um / (hm - hr)
    xi = (x - x0) / t
    return np.where(xi < -cl, hl, np.where(xi <= um - cm, (2 * cl - xi) ** 2 / (9 * G),
                                           np.where(xi < s, hm, hr)))


@pytest.mark.parametrize("mode,tol", [(O.TEXTBOOK, 0.004), (O.FAITHFUL, 0.02)])
def test_stoker_dam_break(mode, tol):
    from rift_amd import synth
    cfg = synth.dam_break(256)
    o = O.Oracle(cfg["dem"], 1.0, h0=cfg["h0"], mode=mode)
    while o.time[0] < 20.0:
        o.step()
    t = o.time[0]
    h = o.interior(o.depth())
    xc = np.arange(h.shape[1]) + 0.5
    band = h[100:156].mean(axis=0)      # rows the y-boundary rarefactions miss
    ex = _stoker(xc, t, 127.5, 1.0, 0.1)
    sel = (xc > 75) & (xc < 225)
    err = np.abs(band[sel] - ex[sel]).mean()
    assert err < tol, err


# --- tiles --------------------------------------------------------------------

def _grow_reference_rule(nbx, nby, wet):
    out = np.empty(nbx * nby, dtype=np.int64)
    for by in range(nby):
        for bx in range(nbx):
            t = by * nbx + bx
            nb = [wet[t]]
            if bx > 0: nb.append(wet[t - 1])
            if bx < nbx - 1: nb.append(wet[t + 1])
            if by > 0: nb.append(wet[t - nbx])
            if by < nby - 1: nb.append(wet[t + nbx])
            out[t] = t if any(nb) else -t
    return out


@pytest.mark.parametrize("nbx,nby", [(3, 3), (4, 3), (2, 5), (1, 4)])
def test_grow_brute_force_all_patterns(nbx, nby):
    L = O.lib()
    n = nbx * nby
    out = np.empty(n, dtype=np.int32)
    for bits in range(1 << n):
        wet = np.array([(bits >> k) & 1 for k in range(n)], dtype=np.uint8)
        L.orc_grow_tiles(nbx, nby, wet.ctypes.data_as(O._U8), out.ctypes.data_as(O._I))
        want = _grow_reference_rule(nbx, nby, wet)
        np.testing.assert_array_equal(out, want)
        # tile 0 is never excluded from the list (-0 == 0)
        active = {int(v) for v in out if v > -1}
        assert 0 in active


def test_timestep_formula():
    # one flux on a still lake of depth h: speeds sqrt(g h); dt as the host
    # computes it: dx / fmaxf(max/C, kappa)
    dem = _flat(32)
    h0 = np.full_like(dem, 2.0)
    o = O.Oracle(dem, 2.0, h0=h0, courant=0.3, mode=O.FAITHFUL, t0=1.0)
    o.phase("grow")
    o.phase("compute_fluxes")
    o.phase("compute_timestep")
    t, dt = o.time
    c = float(np.float32(math.sqrt(np.float32(G * 2.0))))
    kap = O.lib().orc_kappa(2.0, 2.0)
    want = 2.0 / float(np.float32(max(np.float32(c / 0.3), np.float32(kap))))
    assert dt == want
    assert t == 1.0 + want


def test_first_step_dt_and_initial_dt():
    o = O.Oracle(_flat(16), 1.0, h0=np.full((17, 17), 0.5))
    t, dt = o.time
    assert dt == float(np.float32(1e-7))
    o.step()
    t, dt = o.time
    assert dt == float(np.float32(1e-6)) and t == dt


def test_fp32_baseline_tracks_fp64():
    from rift_amd import synth
    cfg = synth.dam_break(64)
    a = O.Oracle(cfg["dem"], 1.0, h0=cfg["h0"], mode=O.FAITHFUL)
    b = O.Oracle(cfg["dem"], 1.0, h0=cfg["h0"], mode=O.FAITHFUL, fp32=True)
    a.step(40)
    b.step(40)
    assert abs(a.time[0] - b.time[0]) < 1e-5 * a.time[0]
    assert np.abs(a.interior(a.depth()) - b.interior(b.depth())).max() < 1e-4


# --- FP64 volume budget ------------------------------------------------------

def _budget_case(flags, surge=False, n=48):
    N = n + 1
    dem = 6.0 * synth.fractal_dem(n, 0.6, 21) + np.linspace(0, 3, N)[None, :]
    h0 = np.maximum(3.5 - dem, 0.0)
    K = 5.0 * np.exp(0.5 * synth.normal(5, N * N).reshape(N, N))
    o = O.Oracle(dem, 2.0, h0=h0, n_const=0.03, K=K, psi=0.11, dtheta=0.3, flags=flags)
    o.set_series_hours(O.SER_HYETO, np.array([0, 0.01, 0.05], np.float32),
                       np.array([50, 120, 0], np.float32))
    o.set_series_hours(O.SER_DRAIN, np.array([0, 0.05], np.float32), np.array([10, 10], np.float32))
    o.set_sources([j * o.nx + i for i in range(20, 23) for j in range(30, 32)])
    o.set_series_hours(O.SER_HYDRO, np.array([0, 0.01], np.float32), np.array([20, 30], np.float32))
    if surge:
        maps = np.full((4, N, N), -9999.0)
        for k in range(4):
            maps[k, :, :3] = 0.5 + 0.4 * k     # surge depth on the west edge
        o.set_map_series(O.MAP_SURGE, maps, 10, 40.0)
    return o


@pytest.mark.parametrize("case", ["heun", "surge", "euler", "no_accel"])
def test_oracle_volume_budget_closes(case):
    """The change of stored + infiltrated volume equals rain + source - drain
    - outflow + exchange to FP64 rounding; each term is exercised."""
    base = O.RAIN_AVG | O.DRAIN_AVG | O.HYDROGRAPH
    flags = {"heun": O.TILE_ACCEL | base, "surge": O.TILE_ACCEL | base | O.SURGE_GRID,
             "euler": O.TILE_ACCEL | base | O.EULER, "no_accel": base}[case]
    o = _budget_case(flags, surge=(case == "surge"))
    for _ in range(4):
        o.step(40)
        b = o.budget()
        scale = np.abs(b[:8]).max()
        assert abs(b[8]) <= 1e-12 * scale, (case, b)
    stored, infil, v0, rain, src, drain, out, exch, _ = b
    assert rain > 0 and src > 0 and drain > 0 and out > 0 and infil > 0
    if case == "surge":
        assert abs(exch) > 1e-3 * scale
    # a dropped term would not close: the outflow alone is far above rounding
    assert out > 1e6 * abs(b[8])
