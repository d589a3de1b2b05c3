"""The surrogate configurations at their stated sizes, on the GPU.

  * warm start of 1e6 ragged columns (layers U{20..40}, 5 % of the inputs
    outside the training ranges): prediction and interpolation to each
    column's layers on the device (sedsurr_predict_device,
    sedcol_set_state_profiles; 2000 sampled columns against the host path
    within 1e-12 relative), one simulated hour of integration with retirement by
    the steady-state residual; every concentration finite and non-negative,
    the C/N/P budget of every column closed to 1e-12, and the retirement of
    each column consistent with its criterion (retired exactly when its
    final residual is below the tolerance);
  * training at batch 4096, where the weight gradients run as 8 k-slices:
    loss and gradient against the network's CPU oracle, and two Adam steps.
"""
import time
from pathlib import Path

import numpy as np
import pytest

import mlp_oracle as M
from sedcol import capi, spinup, surrogate
from helpers import budget_residuals

pytestmark = pytest.mark.gpu

MODEL = Path(surrogate.__file__).resolve().parent / "models" / "spinup_mlp.npz"


def test_warm_start_one_million_columns(gpu_ready):
    t0 = time.perf_counter()
    n = 1_000_000
    rng = np.random.Generator(np.random.Philox(21))
    s = spinup.draw_inputs(rng, n, tails=0.05)
    K = rng.integers(20, 41, size=n)
    dom = spinup.build(s, K)
    print(f"built {n} columns, {dom.ncell} cells ({time.perf_counter() - t0:.1f} s)", flush=True)
    model = surrogate.Model.load(MODEL)
    Xn = (spinup.as_matrix(s) - model.x_mean) / model.x_std
    sample = np.sort(rng.choice(n, 2000, replace=False))
    tol = 1.0e-6
    with surrogate.Network(model.shape, max_batch=4096) as net, capi.SedimentColumns.from_domain(dom) as g:
        net.set_params(model.params)
        # host path on a sample of rows, for the comparison below
        yn_host = net.predict(Xn[sample])
        yn = net.predict_device(Xn)
        assert (yn.rows, yn.ld) == (n, 330)
        t1 = time.perf_counter()
        surrogate.warm_init(model, yn, s["depth"])(g)
        print(f"predicted on the device; state interpolated on the device in {time.perf_counter() - t1:.3f} s "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
        x0 = g.get_state()
        assert np.all(np.isfinite(x0)) and np.all(x0 >= 0.0)
        # the sampled columns against the host interpolation (the node grids
        # may differ in the last bit: pow of the device and of the host)
        off = dom.off
        cells = np.concatenate([np.arange(off[c], off[c + 1]) for c in sample])
        ref = spinup.from_targets(yn_host.astype(np.float64) * model.y_std + model.y_mean, s["depth"][sample],
                                  K[sample])
        assert np.all(np.abs(x0[:, cells] - ref) <= 1e-12 * np.abs(ref) + 1e-14)
        del x0
        inv0 = g.inventory()
        tc, rs, nu = g.spinup(0.0, 3600.0, 3600.0, tol, spinup.CFLOOR)
        x = g.get_state()
        inv1 = g.inventory()
        b = g.budget()
        d = g.diagnostics()
    print(f"1 h of integration: {int((tc >= 0).sum())} columns retired, {nu} not "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    assert np.all(np.isfinite(x)) and np.all(x >= 0.0)
    assert np.all(d["status"] == 0)
    assert np.array_equal(tc >= 0.0, rs < tol) and nu == int((tc < 0.0).sum())
    res = budget_residuals(inv0, inv1, b)
    for e, r in res.items():
        print(f"budget {e}: max rel closure {r.max():.3e}", flush=True)
        assert r.max() <= 1e-12


def test_batch_4096_gradient_and_adam_match_oracle(gpu_ready):
    rng = np.random.default_rng(11)
    shape = surrogate.SHAPE
    p = surrogate.init_params(rng, shape)
    X = rng.standard_normal((4096, 12)).astype(np.float32)
    Y = rng.standard_normal((4096, 330)).astype(np.float32)
    with surrogate.Network(shape, max_batch=4096) as net:
        net.set_params(p)
        lg, gg = net.loss_grad(X, Y)
        lo, go, _ = M.loss_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), shape)
        print(f"batch 4096: loss gpu {lg:.7e} oracle {lo:.7e}", flush=True)
        assert abs(lg - lo) <= 1e-4 * lo
        off = 0
        for (Wo, bo) in M.unflatten(go, *shape):
            for ref in (Wo.ravel(), bo):
                got = gg[off:off + ref.size].astype(np.float64)
                off += ref.size
                err = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30))
                print(f"  tensor of {ref.size}: rel L2 error {err:.2e}", flush=True)
                assert err <= 2e-2
        # two Adam steps on two batches of 4096 rows
        Xs = rng.standard_normal((8192, 12)).astype(np.float32)
        Ys = rng.standard_normal((8192, 330)).astype(np.float32)
        net.set_params(p)
        net.set_data(Xs, Ys)
        rows = rng.permutation(8192).astype(np.int32)
        loss_g = net.train(rows, 4096, lr=1e-3)
        wg = net.get_params().astype(np.float64)
    w = p.astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    loss_o = []
    for st in range(2):
        r = rows[st * 4096:(st + 1) * 4096]
        lo, g, _ = M.loss_grad(w.astype(np.float32).astype(np.float64), Xs[r].astype(np.float64),
                               Ys[r].astype(np.float64), shape)
        loss_o.append(lo)
        w, m, v = M.adam_step(w, g, m, v, st + 1, 1e-3, 0.9, 0.999, 1e-8)
    print("Adam losses gpu", loss_g, "oracle", np.array(loss_o), flush=True)
    assert np.allclose(loss_g, loss_o, rtol=1e-3)
    dg, do = wg - p, w - p
    corr = float(dg @ do / (np.linalg.norm(dg) * np.linalg.norm(do)))
    print("update correlation", corr, flush=True)
    assert corr >= 0.99
