"""U-Net flood surrogate (include/rift_unet.h) and ensemble screening.

`UNet` wraps a rift_unet handle; its inputs and outputs live in device
memory allocated through the same library (kp.DeviceArray), so that solver
maps reach the training set without a host round trip.  The screening workflow (`screen`) predicts the maximum-depth and
arrival-time maps of many scenarios, decides each scenario on a threshold of
its predicted depth at an asset, and re-runs with the solver every scenario
whose prediction lies within a margin of the threshold, plus a held-out share
that measures the surrogate's error.  The reference has no learned component;
the labels are the solver's own `save_max` / `save_arrival` maps
(peak_flood_depth / flood_wave_arrival, grid.cu:930-966).
"""
import ctypes as C

import numpy as np

from . import kp
from .kp import _check, load


def _addr(x):
    """Device address of a DeviceArray (element 0) or an integer pointer."""
    return C.c_void_p(x.ptr() if isinstance(x, kp.DeviceArray) else int(x))


class UNet:
    """A rift_unet handle: `levels` levels of base << l channels."""

    def __init__(self, height=256, width=256, in_channels=4, out_channels=2, base=32, levels=4,
                 max_batch=16, seed=1, device=0):
        L = load()
        self.desc = kp.UNetDesc(height, width, in_channels, out_channels, base, levels, max_batch,
                                seed, device)
        h = C.c_void_p()
        _check(L.rift_unet_create(C.byref(self.desc), C.byref(h)), "rift_unet_create")
        self.h = h
        self.height, self.width = height, width
        self.in_channels, self.out_channels = in_channels, out_channels
        self.max_batch = max_batch
        n = C.c_long()
        _check(L.rift_unet_param_count(self.h, C.byref(n)), "rift_unet_param_count")
        self.param_count = n.value

    def get_params(self):
        p = np.empty(self.param_count, np.float32)
        _check(load().rift_unet_get_params(self.h, p.ctypes.data_as(kp._F)), "rift_unet_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.param_count:
            raise kp.RiftError(f"set_params: expected {self.param_count} values")
        _check(load().rift_unet_set_params(self.h, p.ctypes.data_as(kp._F)), "rift_unet_set_params")

    def set_loss_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != self.out_channels:
            raise kp.RiftError("set_loss_weights: one weight per output channel")
        _check(load().rift_unet_set_loss_weights(self.h, w.ctypes.data_as(kp._F)), "rift_unet_set_loss_weights")

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(load().rift_unet_set_adam(self.h, beta1, beta2, eps), "rift_unet_set_adam")

    def _batch(self, batch):
        if not 1 <= batch <= self.max_batch:
            raise kp.RiftError(f"batch must lie in [1, {self.max_batch}]")
        return int(batch)

    def forward(self, x, batch, y):
        """y = net(x) on device arrays (DeviceArray or address) of
        [batch][in_channels][H][W] and [batch][out_channels][H][W] floats."""
        _check(load().rift_unet_forward(self.h, _addr(x), self._batch(batch), _addr(y)), "rift_unet_forward")

    def gradient(self, x, target, batch):
        g = np.empty(self.param_count, np.float32)
        loss = C.c_double()
        _check(load().rift_unet_gradient(self.h, _addr(x), _addr(target), self._batch(batch), C.byref(loss),
                                         g.ctypes.data_as(kp._F)), "rift_unet_gradient")
        return loss.value, g

    def train_step(self, x, target, batch, lr):
        loss = C.c_double()
        _check(load().rift_unet_train_step(self.h, _addr(x), _addr(target), self._batch(batch), float(lr),
                                           C.byref(loss)), "rift_unet_train_step")
        return loss.value

    # host-array conveniences (one transfer each way)
    def _io(self, batch):
        xs = (batch, self.in_channels, self.height, self.width)
        ys = (batch, self.out_channels, self.height, self.width)
        return xs, ys

    def forward_np(self, x):
        x = np.ascontiguousarray(x, np.float32)
        xs, ys = self._io(x.shape[0])
        if x.shape != xs:
            raise kp.RiftError(f"forward_np: expected shape {xs}")
        dx, dy = kp.DeviceArray(xs, self.desc.device), kp.DeviceArray(ys, self.desc.device)
        dx.upload(x)
        self.forward(dx, xs[0], dy)
        return dy.download()

    def gradient_np(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        xs, ys = self._io(x.shape[0])
        if x.shape != xs or t.shape != ys:
            raise kp.RiftError("gradient_np: bad shapes")
        dx, dt = kp.DeviceArray(xs, self.desc.device), kp.DeviceArray(ys, self.desc.device)
        dx.upload(x)
        dt.upload(t)
        return self.gradient(dx, dt, xs[0])

    def flops(self, batch):
        f, t = C.c_double(), C.c_double()
        _check(load().rift_unet_flops(self.h, int(batch), C.byref(f), C.byref(t)), "rift_unet_flops")
        return f.value, t.value

    def close(self):
        if getattr(self, "h", None):
            load().rift_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- the screening workflow -------------------------------------------------

SIDE = 256              # cells per side of a surrogate scenario (dx = 10 m)
TF = 3600.0             # simulated time, s
ASSET = 8               # decision window: the ASSET x ASSET cells at the centre
THRESHOLD = 0.3         # m of maximum depth at the asset
MARGIN = 0.1            # predictions within THRESHOLD +- MARGIN go to the solver
FEATURES = ("elevation", "relief_9", "relief_33", "slope_x", "slope_y", "manning_n", "rain_peak", "log_K")
TARGETS = ("max_depth_m", "arrival_h")
SOLVER_FLAGS = kp.TILE_ACCEL | kp.RAIN_AVG | kp.SAVE_MAX | kp.SAVE_ARRIVAL


def scenario(seed, side=SIDE, tf=TF):
    """One seeded scenario in the style of the batched ensemble: fractal DEM
    (0-60 m) with a 3 m channel, Manning n ~ U(0.02, 0.08), triangular
    hyetograph over 1 h with a lognormal peak (median 50 mm/hr, sigma_ln 0.5),
    Green-Ampt K lognormal (median 5 mm/hr, sigma_ln 0.8), psi 0.11,
    dtheta 0.3."""
    from . import synth
    u = synth.uniform(seed, 1, 21)[0]
    z = synth.normal(seed, 1, 22)[0]
    dem = synth.carved_channel(60.0 * synth.fractal_dem(side, 0.6, seed), 3.0, seed)
    N = side + 1
    K = 5.0 * np.exp(0.8 * synth.normal(seed, N * N, 13).reshape(N, N))
    return {"seed": seed, "dem": dem, "dx": 10.0, "n_const": float(0.02 + 0.06 * u), "K": K, "psi": 0.11,
            "dtheta": 0.3, "hours": np.array([0.0, 0.5, 1.0], np.float32),
            "mmhr": np.array([0.0, 50.0 * np.exp(0.5 * z), 0.0], np.float32), "tf": tf, "side": side}


def _box_mean(a, w):
    """Mean over a w x w window (edge-replicated) by summed-area table."""
    r = w // 2
    p = np.pad(a, r, mode="edge")
    c = np.zeros((p.shape[0] + 1, p.shape[1] + 1))
    c[1:, 1:] = p.cumsum(0).cumsum(1)
    n0, n1 = a.shape
    return (c[w:w + n0, w:w + n1] - c[:n0, w:w + n1] - c[w:w + n0, :n1] + c[:n0, :n1]) / (w * w)


def features(sc):
    """Network input (8, side, side) float32 at cell centres (the mean of a
    cell's four corner points, as the solver forms cell values).  The
    network computes in bfloat16 (8 significant bits), so elevation enters
    at the scales flow depends on rather than only as an absolute level:
    elevation above the lowest cell / 60 m; relief against the 9 x 9 and
    33 x 33 cell means / 2 m and / 5 m; central-difference slopes / 0.05;
    (n - 0.05) / 0.02; ln(peak / 50 mm/hr) / 0.5; ln(K / 5 mm/hr) / 0.8."""
    def centre(a):
        return 0.25 * (a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:])
    z = centre(sc["dem"])
    dx = sc["dx"]
    zp = np.pad(z, 1, mode="edge")
    out = np.empty((len(FEATURES),) + z.shape, np.float32)
    out[0] = (z - z.min()) / 60.0
    out[1] = (z - _box_mean(z, 9)) / 2.0
    out[2] = (z - _box_mean(z, 33)) / 5.0
    out[3] = (zp[1:-1, 2:] - zp[1:-1, :-2]) / (2.0 * dx) / 0.05
    out[4] = (zp[2:, 1:-1] - zp[:-2, 1:-1]) / (2.0 * dx) / 0.05
    out[5] = (sc["n_const"] - 0.05) / 0.02
    out[6] = np.log(float(sc["mmhr"].max()) / 50.0) / 0.5
    out[7] = centre(np.log(sc["K"] / 5.0) / 0.8)
    return out


def _gen(args):
    seed, side, tf = args
    sc = scenario(seed, side, tf)
    return sc, features(sc)


def generate(seeds, side=SIDE, tf=TF, workers=1, start="fork"):
    """(scenarios, features) for the seeds; `workers` > 1 uses a process pool
    (forked: create it before the process touches the GPU; start="spawn"
    starts fresh interpreters, for a process that already has)."""
    args = [(int(s), side, tf) for s in seeds]
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context(start).Pool(workers) as pool:
            res = pool.map(_gen, args, chunksize=8)
    else:
        res = [_gen(a) for a in args]
    return [r[0] for r in res], np.stack([r[1] for r in res])


DT_FLOOR = 1e-3         # s; a member whose dt falls below it has blown up


def _run_guarded(e, part, log, t_start, check_every=2048, max_steps=150000):
    """Steps the ensemble until every member has ended.  A member whose dt
    falls below DT_FLOOR (or whose time stops being finite) is ended where it
    is and reported as failed: the scheme itself can blow up (seed 10441 of
    the test set does at t = 1569.5 s in the FP64 oracle exactly as on the
    device), and one such member would otherwise hold the whole batch at a
    vanishing dt.  Members still running after max_steps are failed too: an
    hour in 150,000 steps is a mean dt of 24 ms, wave speeds above 100 m/s
    at dx = 10 m, which only a scheme gone astray reaches, and such a member
    would hold its ensemble, alone, for as long as it runs (in a logged run
    every other member of a batch had ended by 70,000 steps)."""
    import time
    steps, failed = 0, set()
    last = time.perf_counter()
    while True:
        e.step(check_every)
        steps += check_every
        times = [g.time for g in e.members]
        for i, ((t, dt), s) in enumerate(zip(times, part)):
            if i in failed or t >= s["tf"]:
                continue
            if not (dt >= DT_FLOOR) or not np.isfinite(t):
                failed.add(i)
                e.set_tf(i, t if np.isfinite(t) and t > 0 else 1e-300)
        if e.finished() == len(part):
            return steps, sorted(failed)
        if steps >= max_steps:
            failed |= {i for i, ((t, _), s) in enumerate(zip(times, part)) if not t >= s["tf"]}
            return steps, sorted(failed)
        if log and time.perf_counter() - last > 15.0:
            last = time.perf_counter()
            log(f"  ... {steps} steps, {e.finished()}/{len(part)} members ended "
                f"({last - t_start:.0f} s)")


def solve(scs, chunk=256, device=0, log=None):
    """The plain batched solver on the scenarios, `chunk` members per
    ensemble.  Returns labels (n, 2, side, side) float32 -- maximum depth (m)
    and arrival time (h; never wet = tf) -- and timings: setup_s (grid
    creation, geometry, uploads), run_s (the stepping), copy_s.  The labels
    of members that blew up (stats["failed"]) are NaN."""
    import time
    n = len(scs)
    side = scs[0]["side"]
    hw = side * side
    labels = np.empty((n, 2, side, side), np.float32)
    stats = {"setup_s": 0.0, "run_s": 0.0, "copy_s": 0.0, "steps": 0, "failed": []}
    for c0 in range(0, n, chunk):
        part = scs[c0:c0 + chunk]
        t0 = time.perf_counter()
        spec = [dict(dem=s["dem"], dx=s["dx"], n_const=s["n_const"], K=s["K"], psi=s["psi"],
                     dtheta=s["dtheta"], flags=SOLVER_FLAGS, device=device) for s in part]
        e = kp.Ensemble(spec)
        from . import synth
        for i, (g, s) in enumerate(zip(e.members, part)):
            t, v = synth.series_hours_to_seconds(s["hours"], s["mmhr"])
            g.set_series(kp.SER_HYETO, t, v)
            e.set_tf(i, s["tf"])
        e.members[0].sync()
        t1 = time.perf_counter()
        steps, failed = _run_guarded(e, part, log, t1)
        stats["steps"] = max(stats["steps"], steps)
        e.members[0].sync()
        t2 = time.perf_counter()
        lab = kp.DeviceArray((len(part), 2, side, side), device)
        for i, g in enumerate(e.members):
            g.copy_field("h_max", lab.ptr(2 * i * hw), 2, 2, side, side)
            g.copy_field("t_wet", lab.ptr((2 * i + 1) * hw), 2, 2, side, side)
        e.members[0].sync()
        arr = lab.download()
        lab.free()
        e.close()
        t3 = time.perf_counter()
        tf = np.array([s["tf"] for s in part], np.float32)[:, None, None]
        arr[:, 1] = np.where(arr[:, 1] < 0.0, tf, arr[:, 1]) / 3600.0
        arr[failed] = np.nan
        labels[c0:c0 + len(part)] = arr
        stats["failed"] += [c0 + i for i in failed]
        stats["setup_s"] += t1 - t0
        stats["run_s"] += t2 - t1
        stats["copy_s"] += t3 - t2
        if log:
            log(f"solved {c0 + len(part)}/{n} (setup {t1 - t0:.1f} s, run {t2 - t1:.1f} s)")
    return labels, stats


def asset_depth(maps):
    """Decision quantity per scenario: the largest maximum depth over the
    ASSET x ASSET cells at the centre of the domain.  maps (n, side, side)."""
    side = maps.shape[-1]
    a = side // 2 - ASSET // 2
    return maps[:, a:a + ASSET, a:a + ASSET].reshape(len(maps), -1).max(axis=1)


def train(net, X, Y, epochs, batch, lr, seed=0, log=None, lr_end_frac=0.05):
    """Adam over the (n, 4, H, W) inputs X and (n, 2, H, W) labels Y held in
    device memory, `epochs` passes in seeded random order, cosine learning
    rate from lr to lr * lr_end_frac.  Returns per-epoch mean losses and the
    training seconds."""
    import time
    n = X.shape[0]
    sx = int(np.prod(X.shape[1:]))   # floats per sample
    sy = int(np.prod(Y.shape[1:]))
    dX = kp.DeviceArray(X.shape, net.desc.device)
    dY = kp.DeviceArray(Y.shape, net.desc.device)
    dX.upload(X)
    dY.upload(Y)
    bX = kp.DeviceArray((batch,) + X.shape[1:], net.desc.device)
    bY = kp.DeviceArray((batch,) + Y.shape[1:], net.desc.device)
    rng = np.random.default_rng(seed)
    steps_per_epoch = n // batch
    total = max(1, epochs * steps_per_epoch)
    hist = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(n)
        acc = 0.0
        for s in range(steps_per_epoch):
            idx = order[s * batch:(s + 1) * batch]
            for j, k in enumerate(idx):
                bX.copy_from(dX, sx, j * sx, int(k) * sx)
                bY.copy_from(dY, sy, j * sy, int(k) * sy)
            f = step / total
            cur = lr * (lr_end_frac + (1.0 - lr_end_frac) * 0.5 * (1.0 + np.cos(np.pi * f)))
            acc += net.train_step(bX, bY, batch, cur)
            step += 1
        hist.append(acc / max(steps_per_epoch, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    seconds = time.perf_counter() - t0
    for a in (dX, dY, bX, bY):
        a.free()
    return hist, seconds


def predict(net, X, batch=None):
    """Network outputs for inputs X (n, 4, H, W) (host), in batches; returns
    (n, 2, H, W) float32 and the seconds spent (uploads, network, downloads)."""
    import time
    batch = batch or net.max_batch
    n = X.shape[0]
    out = np.empty((n, net.out_channels, net.height, net.width), np.float32)
    bX = kp.DeviceArray((batch,) + X.shape[1:], net.desc.device)
    bY = kp.DeviceArray((batch,) + out.shape[1:], net.desc.device)
    t0 = time.perf_counter()
    for s in range(0, n, batch):
        m = min(batch, n - s)
        bX.upload(X[s:s + m])
        net.forward(bX, m, bY)
        out[s:s + m] = bY.download(m * int(np.prod(out.shape[1:]))).reshape((m,) + out.shape[1:])
    seconds = time.perf_counter() - t0
    bX.free()
    bY.free()
    return out, seconds


def error_stats(pred, truth):
    """Maximum-depth and arrival errors of predictions against solver maps
    (n, 2, H, W): mean absolute error, RMSE, the 95th percentile of the
    per-scenario mean absolute error, and the asset-depth error."""
    dh = np.abs(pred[:, 0] - truth[:, 0])
    per = dh.reshape(len(dh), -1).mean(axis=1)
    qa, qt = asset_depth(pred[:, 0]), asset_depth(truth[:, 0])
    return {"depth_mae_m": float(dh.mean()), "depth_rmse_m": float(np.sqrt((dh ** 2).mean())),
            "depth_mae_p95_m": float(np.percentile(per, 95)),
            "asset_depth_mae_m": float(np.abs(qa - qt).mean()),
            "arrival_mae_h": float(np.abs(pred[:, 1] - truth[:, 1]).mean())}


INFER_BATCH = 64


def screen(net, scs, X, heldout_frac=0.05, seed=1, chunk=256, log=None):
    """Surrogate screening of the scenarios (features X): predict every
    scenario, send to the solver those whose predicted asset depth lies
    within MARGIN of THRESHOLD plus a seeded held-out share, and measure the
    surrogate's error on the held-out scenarios.  Returns the decisions,
    the final maps (solver maps where the solver ran) and the timings."""
    pred, t_inf = predict(net, X, min(net.max_batch, INFER_BATCH))
    q = asset_depth(pred[:, 0])
    near = np.abs(q - THRESHOLD) < MARGIN
    rng = np.random.default_rng(seed)
    held = np.zeros(len(scs), bool)
    held[rng.choice(len(scs), max(1, int(round(heldout_frac * len(scs)))), replace=False)] = True
    run = np.flatnonzero(near | held)
    truth, st = solve([scs[i] for i in run], chunk=chunk, device=net.desc.device, log=log)
    ok = ~np.isnan(truth).any(axis=(1, 2, 3))      # solver runs that did not blow up
    final = pred.copy()
    final[run[ok]] = truth[ok]
    solved = np.zeros(len(scs), bool)
    solved[run[ok]] = True
    pos = {i: k for k, i in enumerate(run)}
    hk = np.array([pos[i] for i in np.flatnonzero(held & solved)], dtype=int)
    err = error_stats(pred[held & solved], truth[hk]) if len(hk) else None
    decided = np.where(near & solved, asset_depth(final[:, 0]) > THRESHOLD, q > THRESHOLD)
    return {"pred": pred, "final": final, "flooded": decided, "fallback": near, "heldout": held,
            "solver_runs": run, "solved": solved, "truth": truth, "infer_s": t_inf, "solver": st,
            "errors": err}


def prepare(n_train, n_test, train_seed0=5000, test_seed0=9000, workers=16, side=SIDE, tf=TF,
            start="fork"):
    """Training and test scenarios with their features (host only)."""
    scs_tr, Xtr = generate(range(train_seed0, train_seed0 + n_train), side, tf, workers, start)
    scs_te, Xte = generate(range(test_seed0, test_seed0 + n_test), side, tf, workers, start)
    return scs_tr, Xtr, scs_te, Xte


def run(n_train=4000, n_test=2048, epochs=20, batch=16, lr=1e-3, train_seed0=5000, test_seed0=9000,
        heldout=0.05, workers=16, chunk=256, solve_all=True, log=None, device=0, data=None):
    """The surrogate-screening workload: solver runs for n_train training
    scenarios (seeds train_seed0 + k), training, screening of n_test new
    scenarios (seeds test_seed0 + k) with solver fallback and a held-out
    share, and -- with solve_all -- the plain batched solver over every test
    scenario as the baseline rate and for the error over all of them.
    Scenario inputs are generated on the host before the timed parts and are
    not counted in either rate; `data` = prepare(...) passes them in when they
    were generated earlier (before the process touched the GPU)."""
    import time
    log = log or (lambda m: None)
    t0 = time.perf_counter()
    if data is None:
        data = prepare(n_train, n_test, train_seed0, test_seed0, workers)
    scs_tr, Xtr, scs_te, Xte = data
    n_train, n_test = len(scs_tr), len(scs_te)
    del data
    gen_s = time.perf_counter() - t0
    log(f"generated {n_train} + {n_test} scenarios in {gen_s:.1f} s")
    Ytr, st_tr = solve(scs_tr, chunk=chunk, device=device, log=log)
    del scs_tr
    keep = ~np.isnan(Ytr).any(axis=(1, 2, 3))      # drop runs that blew up
    Xtr, Ytr = Xtr[keep], Ytr[keep]
    side = scs_te[0]["side"]
    # trained at `batch`, run for inference at batch 64 (more blocks for the
    # small levels)
    net = UNet(side, side, len(FEATURES), 2, 32, 4, max_batch=max(batch, INFER_BATCH), seed=1, device=device)
    net.set_loss_weights([1.0, 0.25])
    max_b = net.max_batch
    fwd_flops, _ = net.flops(max_b)
    _, train_flops = net.flops(batch)
    hist, train_s = train(net, Xtr, Ytr, epochs, batch, lr, seed=2, log=log)
    del Xtr, Ytr
    res = screen(net, scs_te, Xte, heldout_frac=heldout, seed=3, chunk=chunk, log=log)
    st = res["solver"]
    runs = len(res["solver_runs"])
    screened_s = res["infer_s"] + st["setup_s"] + st["run_s"] + st["copy_s"]
    out = {
        "n_train": n_train, "n_test": n_test, "epochs": epochs, "batch": batch, "lr": lr,
        "side": side, "threshold_m": THRESHOLD, "margin_m": MARGIN, "asset_cells": ASSET * ASSET,
        "heldout_frac": heldout, "generate_s": gen_s,
        "train_solver_s": st_tr["setup_s"] + st_tr["run_s"] + st_tr["copy_s"],
        "train_s": train_s, "train_loss": hist,
        "train_tflops": train_flops * epochs * (n_train // batch) / max(train_s, 1e-9) / 1e12,
        "infer_s": res["infer_s"], "infer_tflops": fwd_flops / max_b * n_test / max(res["infer_s"], 1e-9) / 1e12,
        "train_solver_failed": len(st_tr["failed"]), "screen_solver_failed": len(st["failed"]),
        "fallback": int(res["fallback"].sum()), "heldout": int(res["heldout"].sum()), "solver_runs": runs,
        "screened_s": screened_s, "screened_per_hour": n_test / screened_s * 3600.0,
        "heldout_errors": res["errors"],
    }
    if solve_all:
        Yte, st_all = solve(scs_te, chunk=chunk, device=device, log=log)
        plain_s = st_all["setup_s"] + st_all["run_s"] + st_all["copy_s"]
        ok = ~np.isnan(Yte).any(axis=(1, 2, 3))
        truth_flood = asset_depth(Yte[ok, 0]) > THRESHOLD
        out.update({
            "plain_s": plain_s, "plain_per_hour": n_test / plain_s * 3600.0,
            "plain_solver_failed": int((~ok).sum()),
            "gain": (n_test / screened_s) / (n_test / plain_s),
            "all_errors_surrogate": error_stats(res["pred"][ok], Yte[ok]),
            "all_errors_screened": error_stats(res["final"][ok], Yte[ok]),
            "decision_accuracy": float(np.mean(res["flooded"][ok] == truth_flood)),
            "flooded_frac": float(truth_flood.mean()),
        })
        # the fallback runs alone (held-out runs as a validation cost apart)
        fb_only = int((res["fallback"] & ~res["heldout"]).sum())
        est = res["infer_s"] + (st["setup_s"] + st["run_s"] + st["copy_s"]) * fb_only / max(runs, 1)
        out["gain_without_heldout_est"] = (n_test / est) / (n_test / plain_s)
    net.close()
    return out


BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 matrix peak


def step_timing(batch=16, steps=10, side=SIDE, device=0):
    """Training-step and forward-pass time of the network at `batch` on
    random inputs (device resident): ms per step and TFLOP/s."""
    import time
    net = UNet(side, side, len(FEATURES), 2, 32, 4, max_batch=batch, seed=1, device=device)
    rng = np.random.default_rng(0)
    x = kp.DeviceArray((batch, len(FEATURES), side, side), device)
    x.upload(rng.standard_normal(x.shape).astype(np.float32))
    t = kp.DeviceArray((batch, 2, side, side), device)
    t.upload(rng.random(t.shape).astype(np.float32))
    y = kp.DeviceArray((batch, 2, side, side), device)
    for _ in range(3):
        net.train_step(x, t, batch, 1e-4)
    t0 = time.perf_counter()
    for _ in range(steps):
        net.train_step(x, t, batch, 1e-4)
    t1 = time.perf_counter()
    for _ in range(steps):
        net.forward(x, batch, y)
    t2 = time.perf_counter()
    f, tr = net.flops(batch)
    net.close()
    ms_tr, ms_f = 1e3 * (t1 - t0) / steps, 1e3 * (t2 - t1) / steps
    return {"batch": batch, "side": side, "train_ms_per_step": ms_tr, "train_tflops": tr / ms_tr / 1e9,
            "forward_ms": ms_f, "forward_tflops": f / ms_f / 1e9,
            "forward_gflop_per_scenario": f / batch / 1e9, "train_gflop_per_scenario": tr / batch / 1e9}


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="Surrogate-screening workload on one GPU")
    ap.add_argument("--train", type=int, default=4000)
    ap.add_argument("--test", type=int, default=2048)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=None)
    ap.add_argument("--step-timing", action="store_true",
                    help="only time training steps and forward passes at --batch on random inputs")
    a = ap.parse_args()
    if a.step_timing:
        print(json.dumps(step_timing(a.batch)))
        sys.exit(0)
    r = run(a.train, a.test, a.epochs, a.batch, a.lr, workers=a.workers,
            log=lambda m: print("[surrogate] " + m, file=sys.stderr, flush=True))
    txt = json.dumps(r)
    print(txt)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
