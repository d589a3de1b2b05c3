"""Surrogate kernels (libsedsurr.so) against the CPU oracle of the network.

The oracle rounds to bfloat16 at the points where the kernels store bfloat16
tensors and accumulates in float64 (the MFMA units accumulate in float32), so
the comparisons use tolerances, stated in each test: relative L2 error of
the loss, of every parameter tensor's gradient and of the outputs.  A
transposed operand, a dropped GELU' or a wrong accumulator layout gives
errors of order 1.
"""
import numpy as np
import pytest

import mlp_oracle as M
from sedcol import surrogate

pytestmark = pytest.mark.gpu

SHAPE = surrogate.SHAPE


def _params(seed, shape=SHAPE):
    rng = np.random.default_rng(seed)
    p = surrogate.init_params(rng, shape).astype(np.float64)
    layers = M.unflatten(p, *shape)
    layers = [(W, 0.1 * rng.standard_normal(b.size)) for W, b in layers]
    return M.flatten(layers).astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def net(gpu_ready):
    n = surrogate.Network(SHAPE, max_batch=256)
    yield n
    n.close()


def test_loss_and_gradient_match_oracle(net):
    p = _params(0)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((200, 12)).astype(np.float32)  # 200: a partial 64-row tile
    Y = rng.standard_normal((200, 330)).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    lg, gg = net.loss_grad(X, Y)
    lo, go, _ = M.loss_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), SHAPE)
    print(f"loss gpu {lg:.7e} oracle {lo:.7e}", flush=True)
    assert abs(lg - lo) <= 1e-4 * lo
    off = 0
    for (Wo, bo) in M.unflatten(go, *SHAPE):
        for ref in (Wo.ravel(), bo):
            got = gg[off:off + ref.size].astype(np.float64)
            off += ref.size
            err = _rel(got, ref)
            print(f"  tensor of {ref.size}: rel L2 error {err:.2e}", flush=True)
            assert err <= 2e-2
    cos = float(gg.astype(np.float64) @ go / (np.linalg.norm(gg) * np.linalg.norm(go)))
    assert cos >= 0.9998


def test_split_k_weight_gradient_matches_oracle(gpu_ready):
    """1000 rows: the weight-gradient products run as two k slices summed
    in order (and a partial last tile of rows)."""
    p = _params(7)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((1000, 12)).astype(np.float32)
    Y = rng.standard_normal((1000, 330)).astype(np.float32)
    with surrogate.Network(SHAPE, max_batch=1024) as n:
        n.set_params(p)
        lg, gg = n.loss_grad(X, Y)
    lo, go, _ = M.loss_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), SHAPE)
    assert abs(lg - lo) <= 1e-4 * lo
    off = 0
    for (Wo, bo) in M.unflatten(go, *SHAPE):
        for ref in (Wo.ravel(), bo):
            err = _rel(gg[off:off + ref.size].astype(np.float64), ref)
            off += ref.size
            print(f"  tensor of {ref.size}: rel L2 error {err:.2e}", flush=True)
            assert err <= 2e-2


def test_predict_matches_oracle_in_passes(net):
    p = _params(2)
    net.set_params(p)
    X = np.random.default_rng(3).standard_normal((600, 12)).astype(np.float32)  # 3 passes of 256
    yg = net.predict(X).astype(np.float64)
    yo = M.forward(p.astype(np.float64), X.astype(np.float64), SHAPE)
    err = _rel(yg, yo)
    print("predict rel L2 error", err, flush=True)
    assert err <= 5e-3
    assert np.max(np.abs(yg - yo)) <= 2e-2 * np.max(np.abs(yo))
    ms, launches = net.last_timing()
    assert launches == 3 * (1 + 4) and ms > 0.0


def test_adam_steps_match_oracle(net):
    p = _params(4)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((512, 12)).astype(np.float32)
    Y = rng.standard_normal((512, 330)).astype(np.float32)
    net.set_params(p)
    net.set_data(X, Y)
    b, steps, lr = 128, 3, 1e-3
    rows = rng.integers(0, 512, size=steps * b).astype(np.int32)
    loss_g = net.train(rows, b, lr=lr)
    wg = net.get_params().astype(np.float64)
    w = p.astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    loss_o = []
    for s in range(steps):
        r = rows[s * b:(s + 1) * b]
        lo, g, _ = M.loss_grad(w.astype(np.float32).astype(np.float64), X[r].astype(np.float64),
                               Y[r].astype(np.float64), SHAPE)
        loss_o.append(lo)
        w, m, v = M.adam_step(w, g, m, v, s + 1, lr, 0.9, 0.999, 1e-8)
    print("losses gpu", loss_g, "oracle", np.array(loss_o), flush=True)
    assert np.allclose(loss_g, loss_o, rtol=1e-3)
    dg, do = wg - p, w - p
    corr = float(dg @ do / (np.linalg.norm(dg) * np.linalg.norm(do)))
    print("update correlation", corr, flush=True)
    assert corr >= 0.99
    # an Adam step moves a parameter by at most lr (1 - b1) / sqrt(1 - b2) ~ 3.2 lr
    assert np.max(np.abs(dg)) <= steps * lr * 3.2


def test_training_reduces_loss_on_a_smooth_map(gpu_ready):
    rng = np.random.default_rng(6)
    A = rng.standard_normal((12, 16)) / np.sqrt(12)
    B = rng.standard_normal((16, 330)) / 4.0
    X = rng.standard_normal((6144, 12))
    Y = np.tanh(X @ A) @ B
    model, hist = surrogate.fit(X, Y, epochs=40, batch=512, lr=2e-3, lr_final=2e-4, seed=1)
    print("val loss", hist["val_loss"], flush=True)
    assert hist["val_loss"][-1] <= 0.05 * hist["val_loss"][0]
    Xt = rng.standard_normal((500, 12))
    yp = surrogate.predict(model, Xt)
    yt = np.tanh(Xt @ A) @ B
    assert np.mean((yp - yt) ** 2) <= 0.05 * np.var(yt)


def test_errors_are_status_codes(gpu_ready):
    with pytest.raises(surrogate.SurrogateError) as e:
        surrogate.Network((12, 100, 3, 330))  # width not a multiple of 64
    assert e.value.code == 1
    with surrogate.Network((12, 64, 1, 10), max_batch=64) as n:
        with pytest.raises(surrogate.SurrogateError) as e:
            n.train(np.zeros(64, np.int32), 64)  # no training set
        assert e.value.code == 1
        bad = np.zeros(n.num_params, np.float32)
        bad[3] = np.nan
        with pytest.raises(surrogate.SurrogateError) as e:
            n.set_params(bad)
        assert e.value.code == 2
        n.set_data(np.zeros((10, 12)), np.zeros((10, 10)))
        with pytest.raises(surrogate.SurrogateError) as e:
            n.train(np.full(64, 10, np.int32), 64)  # row outside the set
        assert e.value.code == 1
