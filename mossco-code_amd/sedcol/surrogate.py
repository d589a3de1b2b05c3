"""Spin-up surrogate: binding of libsedsurr.so, training driver, warm start.

The network maps the 12 column inputs of sedcol.spinup (INPUTS) to the
log10 steady-state profiles on 30 layers (330 values); it is trained on
steady states from cold spin-ups (sedcol.spinup.generate) by the library's
own BF16 MFMA kernels, and its prediction, interpolated to each column's
layers, is the warm start that sedcol_spinup continues from.  Columns whose
warm start does not reach the steady-state criterion within the warm budget
fall back to a cold start.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_NAME = "libsedsurr.so"
SHAPE = (12, 512, 3, 330)  # n_in, n_hidden, hidden layers, n_out

#: every symbol include/sedsurr.h declares
EXPORTS = ("sedsurr_create", "sedsurr_destroy", "sedsurr_last_error", "sedsurr_num_params", "sedsurr_set_params",
           "sedsurr_get_params", "sedsurr_set_data", "sedsurr_train", "sedsurr_loss_grad", "sedsurr_predict",
           "sedsurr_predict_device",
           "sedsurr_last_timing")


class Shape(ctypes.Structure):
    _fields_ = [("n_in", ctypes.c_int32), ("n_hidden", ctypes.c_int32), ("n_layers", ctypes.c_int32),
                ("n_out", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Adam(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float)]


class SurrogateError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sedsurr status {code}: {msg}")
        self.code = code


_LIB = None


def lib_path() -> Path:
    return Path(os.environ.get("SEDSURR_LIB", HERE / LIB_NAME))


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not path.exists():
            raise OSError(f"{path} not found: build it with __graft_entry__.build()")
        L = ctypes.CDLL(str(path))
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        pf = ctypes.POINTER(ctypes.c_float)
        sig = {
            "sedsurr_create": (ctypes.c_int, [ctypes.POINTER(Shape), ctypes.POINTER(vp)]),
            "sedsurr_destroy": (ctypes.c_int, [vp]),
            "sedsurr_last_error": (ctypes.c_char_p, []),
            "sedsurr_num_params": (ctypes.c_int, [vp, ctypes.POINTER(i64)]),
            "sedsurr_set_params": (ctypes.c_int, [vp, pf]),
            "sedsurr_get_params": (ctypes.c_int, [vp, pf]),
            "sedsurr_set_data": (ctypes.c_int, [vp, pf, pf, i64]),
            "sedsurr_train": (ctypes.c_int, [vp, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(Adam), pf]),
            "sedsurr_loss_grad": (ctypes.c_int, [vp, pf, pf, i32, pf, pf]),
            "sedsurr_predict": (ctypes.c_int, [vp, pf, i64, pf]),
            "sedsurr_predict_device": (ctypes.c_int, [vp, pf, i64, ctypes.POINTER(vp)]),
            "sedsurr_last_timing": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(code):
    if code != 0:
        msg = lib().sedsurr_last_error()
        raise SurrogateError(code, msg.decode() if msg else "")


def _pf(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Network:
    """A device-resident MLP (libsedsurr.so handle)."""

    def __init__(self, shape=SHAPE, max_batch=4096):
        self.shape = tuple(int(v) for v in shape)
        s = Shape(*self.shape, int(max_batch), 0)
        h = ctypes.c_void_p()
        check(lib().sedsurr_create(ctypes.byref(s), ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        check(lib().sedsurr_num_params(self._h, ctypes.byref(n)))
        self.num_params = n.value
        self.max_batch = int(max_batch)

    def close(self):
        if getattr(self, "_h", None):
            lib().sedsurr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.num_params
        check(lib().sedsurr_set_params(self._h, _pf(p)))

    def get_params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        check(lib().sedsurr_get_params(self._h, _pf(p)))
        return p

    def set_data(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        self._n = X.shape[0]
        check(lib().sedsurr_set_data(self._h, _pf(X), _pf(Y), X.shape[0]))

    def train(self, rows, batch, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        steps = rows.size // batch
        loss = np.empty(steps, dtype=np.float32)
        a = Adam(lr, beta1, beta2, eps)
        check(lib().sedsurr_train(self._h, steps, int(batch), rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  ctypes.byref(a), _pf(loss)))
        return loss

    def loss_grad(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.num_params, dtype=np.float32)
        check(lib().sedsurr_loss_grad(self._h, _pf(X), _pf(Y), X.shape[0], _pf(loss), _pf(g)))
        return float(loss[0]), g

    def predict(self, X):
        X = _f32(X)
        Y = np.empty((X.shape[0], self.shape[3]), dtype=np.float32)
        check(lib().sedsurr_predict(self._h, _pf(X), X.shape[0], _pf(Y)))
        return Y

    def predict_device(self, X):
        """predict with the output left on the GPU, in a buffer of this
        network that stays valid until its next predict_device or close:
        capi.DeviceRows (n rows of n_out), ready for
        SedimentColumns.set_state_profiles."""
        from .capi import DeviceRows

        X = _f32(X)
        ptr = ctypes.c_void_p()
        check(lib().sedsurr_predict_device(self._h, _pf(X), X.shape[0], ctypes.byref(ptr)))
        return DeviceRows(int(ptr.value), int(X.shape[0]), int(self.shape[3]))

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int32()
        check(lib().sedsurr_last_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


def init_params(rng: np.random.Generator, shape=SHAPE) -> np.ndarray:
    """Glorot-uniform weights, zero biases, in the flat layout of
    sedsurr_get_params."""
    n_in, n_hidden, n_layers, n_out = shape
    dims = [n_in] + [n_hidden] * n_layers + [n_out]
    parts = []
    for fi, fo in zip(dims[:-1], dims[1:]):
        a = np.sqrt(6.0 / (fi + fo))
        parts.append(rng.uniform(-a, a, size=fi * fo))
        parts.append(np.zeros(fo))
    return np.concatenate(parts).astype(np.float32)


def flops_per_sample(shape=SHAPE) -> float:
    """Multiply-adds x 2 of one forward pass."""
    n_in, n_hidden, n_layers, n_out = shape
    dims = [n_in] + [n_hidden] * n_layers + [n_out]
    return 2.0 * sum(fi * fo for fi, fo in zip(dims[:-1], dims[1:]))


@dataclass
class Model:
    """Trained parameters and the normalisation of inputs and targets."""
    params: np.ndarray
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray
    shape: tuple = SHAPE

    def save(self, path):
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean,
                 y_std=self.y_std, shape=np.asarray(self.shape))

    @classmethod
    def load(cls, path):
        z = np.load(path)  # plain arrays, no pickles
        return cls(z["params"], z["x_mean"], z["x_std"], z["y_mean"], z["y_std"], tuple(int(v) for v in z["shape"]))


def fit(X, Y, *, epochs=400, batch=4096, lr=1e-3, lr_final=1e-5, val_fraction=0.1, seed=0, log=None,
        net: Network | None = None):
    """Train on (X (n x 12), Y (n x 330)) with Adam on shuffled minibatches,
    learning rate decayed geometrically from lr to lr_final over the epochs.
    Returns (Model, history dict)."""
    import time

    rng = np.random.Generator(np.random.Philox(seed))
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n = X.shape[0]
    perm = rng.permutation(n)
    nval = int(round(val_fraction * n))
    iv, it = perm[:nval], perm[nval:]
    x_mean, x_std = X[it].mean(0), X[it].std(0) + 1e-12
    y_mean, y_std = Y[it].mean(0), Y[it].std(0) + 1e-6
    Xn = (X - x_mean) / x_std
    Yn = (Y - y_mean) / y_std
    own = net is None
    net = net or Network(SHAPE, max_batch=batch)
    try:
        net.set_params(init_params(rng, net.shape))
        net.set_data(Xn[it], Yn[it])
        ntr = it.size
        b = min(batch, ntr - ntr % 64 if ntr >= 64 else ntr)
        steps_per_epoch = max(1, ntr // b)
        hist = {"epoch": [], "train_loss": [], "val_loss": [], "lr": []}
        train_ms = 0.0
        launches = 0
        t0 = time.perf_counter()
        for ep in range(epochs):
            cur = lr * (lr_final / lr) ** (ep / max(1, epochs - 1))
            rows = np.concatenate([rng.permutation(ntr)[: steps_per_epoch * b]])
            loss = net.train(rows, b, lr=cur)
            ms, nl = net.last_timing()
            train_ms += ms
            launches += nl
            if ep % 10 == 0 or ep == epochs - 1:
                vl = float(np.mean((net.predict(Xn[iv]) - Yn[iv]) ** 2)) if nval else float("nan")
                hist["epoch"].append(ep)
                hist["train_loss"].append(float(loss.mean()))
                hist["val_loss"].append(vl)
                hist["lr"].append(cur)
                if log:
                    log(f"epoch {ep} lr {cur:.2e} train {loss.mean():.4e} val {vl:.4e} "
                        f"({time.perf_counter() - t0:.1f} s)")
        hist["train_ms"] = train_ms
        hist["launches"] = launches
        hist["steps"] = epochs * steps_per_epoch
        hist["batch"] = b
        hist["wall_s"] = time.perf_counter() - t0
        model = Model(net.get_params(), x_mean, x_std, y_mean, y_std, net.shape)
    finally:
        if own:
            net.close()
    return model, hist


def warm_init(model: Model, yn_dev, depth, dz1: float = 1.0e-3):
    """init for spinup.spin: the network's normalised prediction yn_dev
    (capi.DeviceRows from Network.predict_device, whose network must stay
    open) denormalised, interpolated to every column's layers and set as its
    state, on the device (sedcol_set_state_profiles).  depth: the columns'
    total depths (m), the depth of the network's 30-layer grid."""
    from .spinup import EPS, K_SURR

    depth = np.asarray(depth, dtype=np.float64)
    return lambda g: g.set_state_profiles(yn_dev, K_SURR, model.y_std, model.y_mean, EPS, dz1, depth)


def predict(model: Model, X, net: Network | None = None):
    """Log10 profiles (n x 330) for raw inputs X (n x 12)."""
    own = net is None
    net = net or Network(model.shape)
    try:
        net.set_params(model.params)
        y = net.predict((np.asarray(X, dtype=np.float64) - model.x_mean) / model.x_std)
    finally:
        if own:
            net.close()
    return y.astype(np.float64) * model.y_std + model.y_mean


def evaluate(model: Model, n: int, seed: int, *, years=2.0, warm_days=30.0, res_tol=1.0e-6, flux_rtol=0.01,
             tails=0.05, adaptive_method=7, dt_max=86400.0, chunk_days=1.0, log=None, dump=None):
    """Surrogate warm start against the plain cold spin-up on n drawn columns
    (layers U{20..40}, a fraction `tails` of the inputs outside the training
    ranges).  Every integration uses the same stepper (adaptive_method,
    steps up to dt_max).

    cold:     the reference's spin-up, a presimulation of `years` from the
              cold state (fabm_sediment_component.F90:557-632), per column;
              its final top fluxes F_cold are the steady-state reference.
    warm:     prediction, interpolation to each column's layers, then the
              integration, a column retiring at the first chunk boundary
              where its residual is below res_tol and every top flux is
              within flux_rtol of F_cold (floor 1 % of the carbon
              deposition); columns not retired within warm_days fall back
              to the cold spin-up (timed and added).
    residual: the same warm integration retiring by the residual alone
              (the criterion a deployment without F_cold would use): the
              fraction retired, and its flux error against F_cold as an
              accuracy figure.
    Returns the wall times of each phase, the ratio cold / warm, the
    fallback rate, and per-variable errors."""
    import time

    from . import capi, spinup

    rng = np.random.Generator(np.random.Philox(seed))
    s = spinup.draw_inputs(rng, n, tails=tails)
    K = rng.integers(20, 41, size=n)
    dom = spinup.build(s, K)
    dom.params.adaptive_method = adaptive_method
    dom.params.emp_dt_max = dt_max
    floor = spinup.flux_floor(s)
    say = log or (lambda m: None)

    say(f"cold: spin-up of {n} columns for {years} y")
    cold = spinup.spin(dom, years, 1e-300, chunk_days=10.0, log=say)
    F_cold = cold["flux"]

    say("warm: prediction and integration")
    X = spinup.as_matrix(s)
    t0 = time.perf_counter()
    # the network holds the prediction on the device until it is closed
    net = Network(model.shape, max_batch=4096)
    try:
        net.set_params(model.params)
        yn = net.predict_device((X - model.x_mean) / model.x_std)
        pred_ms, _ = net.last_timing()
        t_pred = time.perf_counter() - t0
        init = warm_init(model, yn, s["depth"])
        warm = spinup.spin(dom, warm_days / 365.0, res_tol, chunk_days=chunk_days, init=init, flux_ref=F_cold,
                           flux_rtol=flux_rtol, floor=floor, log=say)
        t_interp = warm["init_s"]
        fb = warm["t_conv"] < 0.0
        t_fb = 0.0
        if fb.any():
            say(f"fallback: cold spin-up of {int(fb.sum())} columns")
            r = spinup.spin(dom.subset(np.flatnonzero(fb)), years, 1e-300, chunk_days=10.0, log=say)
            t_fb = r["wall_s"]
        wall_warm = t_pred + t_interp + warm["wall_s"] + t_fb

        say("residual-only retirement (accuracy figure)")
        res_only = spinup.spin(dom, warm_days / 365.0, res_tol, chunk_days=chunk_days, init=init, log=say)
        ro = res_only["t_conv"] >= 0.0
        with capi.SedimentColumns.from_domain(dom) as g:
            init(g)
            x0 = g.get_state()
    finally:
        net.close()

    lg = lambda a: np.log10(np.maximum(a, 0.0) + spinup.EPS[:, None])
    prof_err = np.abs(lg(x0) - lg(cold["state"]))
    sc = np.maximum(np.abs(F_cold), floor[None, :])
    fl_err = np.abs(warm["flux"] - F_cold) / sc
    fl_ro = np.abs(res_only["flux"] - F_cold) / sc
    ok = ~fb
    per_var = {}
    for v, name in enumerate(_VARS):
        per_var[name] = {
            "warm_start_log10_err_median": float(np.median(prof_err[v])),
            "warm_start_log10_err_p95": float(np.quantile(prof_err[v], 0.95)),
            "flux_rel_err_median_retired": float(np.median(fl_err[v][ok])) if ok.any() else None,
            "flux_rel_err_median_residual_only": float(np.median(fl_ro[v][ro])) if ro.any() else None,
            "flux_within_1pct_residual_only": float(np.mean(fl_ro[v][ro] <= 0.01)) if ro.any() else None,
        }
    tail = s["tail"]
    if dump:
        # per-column record: cold attempts and final residual, warm
        # retirement time (-1: fallback), residual-only retirement time
        np.savez_compressed(dump, cold_attempts=cold["attempts"], cold_resid=cold["resid"],
                            warm_t_conv=warm["t_conv"], warm_attempts=warm["attempts"],
                            resonly_t_conv=res_only["t_conv"], tail=tail, nlayers=np.asarray(K),
                            flux_err_warm=fl_err, flux_err_resonly=fl_ro)
    return {
        "columns": n, "tail_fraction": float(tail.mean()), "years": years, "warm_days": warm_days,
        "adaptive_method": adaptive_method, "dt_max_s": dt_max, "res_tol": res_tol, "flux_rtol": flux_rtol,
        "wall_cold_s": cold["wall_s"], "wall_warm_s": wall_warm, "speedup_wall": cold["wall_s"] / wall_warm,
        "wall_predict_s": t_pred, "predict_kernel_ms": pred_ms, "wall_interpolate_s": t_interp,
        "wall_warm_integration_s": warm["wall_s"], "wall_fallback_s": t_fb,
        "fallback_rate": float(fb.mean()),
        "fallback_rate_in_range": float(fb[~tail].mean()) if (~tail).any() else None,
        "fallback_rate_tail": float(fb[tail].mean()) if tail.any() else None,
        "residual_only_retired_fraction": float(ro.mean()),
        "residual_only_all_fluxes_within_1pct": float(np.mean(np.all(fl_ro[:, ro] <= 0.01, axis=0))) if ro.any() else None,
        "warm_days_to_criterion_median": float(np.median(warm["t_conv"][ok]) / 86400.0) if ok.any() else None,
        "cold_attempts_per_column_median": float(np.median(cold["attempts"])),
        "cold_attempts_per_column_max": float(np.max(cold["attempts"])),
        "per_variable": per_var,
    }


_VARS = ("ldetC", "sdetC", "ldetN", "sdetN", "detP", "O2", "NO3", "NH4", "ODU", "PO4", "adsP")


def _main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train the spin-up surrogate / evaluate its warm start")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, nargs="+", help="npz files from python -m sedcol.spinup (X, Y)")
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=600)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-final", type=float, default=2e-5)
    b = sub.add_parser("bench", help="time training steps and prediction on synthetic data of the network's shape")
    b.add_argument("--steps", type=int, default=200)
    b.add_argument("--batch", type=int, default=4096)
    b.add_argument("--rows", type=int, default=1 << 20, help="rows of one prediction call")
    w = sub.add_parser("warm", help="time the device warm start (prediction and interpolation) of n ragged columns")
    w.add_argument("--model", required=True)
    w.add_argument("--n", type=int, default=1_000_000)
    w.add_argument("--seed", type=int, default=21)
    w.add_argument("--repeats", type=int, default=4)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", required=True)
    e.add_argument("--n", type=int, default=10000)
    e.add_argument("--seed", type=int, default=7)
    e.add_argument("--years", type=float, default=2.0)
    e.add_argument("--warm-days", type=float, default=30.0)
    e.add_argument("--method", type=int, default=7, help="adaptive stepper: 6 EMP, 7 linearly implicit ROS2")
    e.add_argument("--dt-max", type=float, default=86400.0)
    e.add_argument("--dump", default=None, help="npz file for the per-column record")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "train":
        zs = [np.load(f) for f in a.data]
        Xa = np.concatenate([z["X"] for z in zs])
        Ya = np.concatenate([z["Y"] for z in zs])
        model, hist = fit(Xa, Ya, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, log=log)
        model.save(a.out)
        n_train = int(round(0.9 * Xa.shape[0]))
        fl = 3.0 * flops_per_sample() * hist["batch"] * hist["steps"]
        print(json.dumps({"train_ms": hist["train_ms"], "steps": hist["steps"], "batch": hist["batch"],
                          "launches": hist["launches"], "final_val_loss": hist["val_loss"][-1],
                          "n_train": n_train, "train_tflops": fl / (hist["train_ms"] * 1e-3) / 1e12}), flush=True)
    elif a.cmd == "bench":
        rng = np.random.Generator(np.random.Philox(0))
        n = 8 * a.batch
        X = rng.standard_normal((n, SHAPE[0])).astype(np.float32)
        Y = rng.standard_normal((n, SHAPE[3])).astype(np.float32)
        with Network(SHAPE, max_batch=a.batch) as net:
            net.set_params(init_params(rng))
            net.set_data(X, Y)
            rows = rng.integers(0, n, size=a.steps * a.batch).astype(np.int32)
            net.train(rows[: 10 * a.batch], a.batch)  # warm-up
            net.train(rows, a.batch)
            ms, launches = net.last_timing()
            fl = 3.0 * flops_per_sample() * a.batch * a.steps
            Xp = rng.standard_normal((a.rows, SHAPE[0])).astype(np.float32)
            net.predict(Xp[: a.batch])
            net.predict(Xp)
            pms, plaunch = net.last_timing()
        print(json.dumps({"train_steps": a.steps, "batch": a.batch, "train_ms": ms, "ms_per_step": ms / a.steps,
                          "launches_per_step": launches / a.steps, "train_tflops": fl / (ms * 1e-3) / 1e12,
                          "predict_rows": a.rows, "predict_ms": pms,
                          "predict_tflops": flops_per_sample() * a.rows / (pms * 1e-3) / 1e12}), flush=True)
    elif a.cmd == "warm":
        import time

        from . import capi, spinup

        model = Model.load(a.model)
        rng = np.random.Generator(np.random.Philox(a.seed))
        s = spinup.draw_inputs(rng, a.n, tails=0.05)
        dom = spinup.build(s, rng.integers(20, 41, size=a.n))
        Xn = (spinup.as_matrix(s) - model.x_mean) / model.x_std
        with Network(model.shape, max_batch=4096) as net, capi.SedimentColumns.from_domain(dom) as g:
            net.set_params(model.params)
            t0 = time.perf_counter()
            yn = net.predict_device(Xn)
            t_pred = time.perf_counter() - t0
            pred_ms, _ = net.last_timing()
            init = warm_init(model, yn, s["depth"])
            walls = []
            for _ in range(a.repeats):
                t0 = time.perf_counter()
                init(g)
                walls.append(time.perf_counter() - t0)
        print(json.dumps({"columns": a.n, "cells": dom.ncell, "predict_s": t_pred, "predict_kernel_ms": pred_ms,
                          "set_state_profiles_s": walls}), flush=True)
    else:
        model = Model.load(a.model)
        r = evaluate(model, a.n, a.seed, years=a.years, warm_days=a.warm_days, adaptive_method=a.method,
                     dt_max=a.dt_max, log=log, dump=a.dump)
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    _main()
