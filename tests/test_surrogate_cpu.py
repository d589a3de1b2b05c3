"""CPU tests of the surrogate's oracle, host logic and C ABI (no GPU).

The oracle (oracle/mlp_oracle.py) is pinned here against library routines
(torch's bfloat16 conversion, GELU and Adam) and against finite differences,
so that the GPU tests can use it as the checker of the kernels.
"""
import ctypes
import re

import numpy as np
import pytest
import torch

import mlp_oracle as M
from conftest import REPO
from sedcol import spinup, surrogate


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.4e38, -3.4e38, 1e-40, np.inf, -np.inf],
                                 dtype=np.float32)])
    # ties: values exactly halfway between two bfloat16 numbers
    u = (rng.integers(0, 1 << 15, 1000, dtype=np.uint32) << 16) | 0x8000
    u = u[((u >> 23) & 0xFF) != 0xFF]  # finite values only
    x = np.concatenate([x, u.view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = M.to_bf16(x)
    assert np.array_equal(got, ref)


def test_gelu_and_derivative_match_torch():
    z = torch.linspace(-12.0, 12.0, 20001, dtype=torch.float64, requires_grad=True)
    g = torch.nn.functional.gelu(z, approximate="tanh")
    g.sum().backward()
    zn = z.detach().numpy()
    assert np.allclose(M.gelu(zn), g.detach().numpy(), rtol=1e-14, atol=1e-15)
    assert np.allclose(M.gelu_grad(zn), z.grad.numpy(), rtol=1e-12, atol=1e-14)


def test_backward_matches_finite_differences():
    shape = (5, 8, 2, 7)
    rng = np.random.default_rng(1)
    p = surrogate.init_params(rng, shape).astype(np.float64) + 0.05 * rng.standard_normal(
        M.num_params(*shape))
    X = rng.standard_normal((9, 5))
    T = rng.standard_normal((9, 7))
    loss, g, _ = M.loss_grad(p, X, T, shape, rounding=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        q = p.copy()
        q[i] += h
        lp = M.loss_grad(q, X, T, shape, rounding=False)[0]
        q[i] -= 2 * h
        lm = M.loss_grad(q, X, T, shape, rounding=False)[0]
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    # a plausible mistake fails: dropping GELU' in the backward pass
    bad = M.gelu_grad
    try:
        M.gelu_grad = lambda z: np.ones_like(z)
        assert not np.allclose(M.loss_grad(p, X, T, shape, rounding=False)[1], fd, atol=1e-6)
    finally:
        M.gelu_grad = bad


def test_adam_matches_torch():
    rng = np.random.default_rng(2)
    w0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(6)]
    wt = torch.tensor(w0, requires_grad=True)
    opt = torch.optim.Adam([wt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(grads, 1):
        wt.grad = torch.tensor(g)
        opt.step()
        w, m, v = M.adam_step(w, g, m, v, t, 1e-3, 0.9, 0.999, 1e-8)
    assert np.allclose(w, wt.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_parameter_layout_and_count():
    shape = surrogate.SHAPE
    n = M.num_params(*shape)
    assert n == 12 * 512 + 512 + 2 * (512 * 512 + 512) + 512 * 330 + 330 == 701258
    p = surrogate.init_params(np.random.default_rng(0), shape)
    assert p.size == n and p.dtype == np.float32
    layers = M.unflatten(p, *shape)
    assert [W.shape for W, _ in layers] == [(12, 512), (512, 512), (512, 512), (512, 330)]
    assert all(np.all(b == 0) for _, b in layers)
    assert np.allclose(M.flatten(layers), p)
    assert surrogate.flops_per_sample() == 2.0 * (12 * 512 + 2 * 512 * 512 + 512 * 330)


def test_header_exports_match_binding():
    hdr = (REPO / "include" / "sedsurr.h").read_text()
    declared = set(re.findall(r"\b(sedsurr_[a-z_]+)\s*\(", hdr))
    assert declared == set(surrogate.EXPORTS)
    L = ctypes.CDLL(str(surrogate.lib_path()))
    for name in surrogate.EXPORTS:
        assert hasattr(L, name), name
    assert ctypes.sizeof(surrogate.Shape) == 24
    assert ctypes.sizeof(surrogate.Adam) == 16


def test_create_without_device_is_a_status():
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(surrogate.SurrogateError) as e:
        surrogate.Network((12, 64, 1, 10), max_batch=64)
    assert e.value.code == 4  # SEDCOL_EHIP


def test_targets_round_trip_and_interpolation():
    rng = np.random.Generator(np.random.Philox(3))
    s = spinup.draw_inputs(rng, 50)
    dom = spinup.build(s, 30)
    x = dom.conc0.copy()
    x[0:5] = rng.uniform(0.0, 5e3, size=x[0:5].shape)
    y = spinup.to_targets(x)
    assert y.shape == (50, 330)
    back = spinup.from_targets(y, s["depth"], 30)
    assert np.allclose(back, x, rtol=1e-12, atol=1e-9)
    # a linear log-profile in depth is reproduced on any layer count
    z30 = np.stack([spinup.layer_centres(30, d) for d in s["depth"]])
    ylin = np.repeat((1.0 + 5.0 * z30)[:, None, :], 11, axis=1).reshape(50, 330)
    for K in (20, 31, 40):
        out = spinup.from_targets(ylin, s["depth"], K)
        zk = np.concatenate([spinup.layer_centres(K, d) for d in s["depth"]])
        inside = np.concatenate([(spinup.layer_centres(K, d) >= z30[i, 0]) & (spinup.layer_centres(K, d) <= z30[i, -1])
                                 for i, d in enumerate(s["depth"])])
        expect = 10.0 ** (1.0 + 5.0 * zk) - spinup.EPS[0]
        assert np.allclose(out[3][inside], expect[inside], rtol=1e-12)


def test_input_draws_and_tails():
    rng = np.random.Generator(np.random.Philox(4))
    s = spinup.draw_inputs(rng, 4000, tails=0.05)
    X = spinup.as_matrix(s)
    assert X.shape == (4000, 12)
    tail = s["tail"]
    assert tail.sum() == 200
    inside = np.ones(4000, dtype=bool)
    for j, k in enumerate(spinup.INPUTS):
        lo, hi = spinup.RANGES[k]
        inside &= (X[:, j] >= lo) & (X[:, j] <= hi)
        llo, lhi = spinup.LIMITS[k]
        assert X[:, j].min() >= llo and X[:, j].max() <= lhi
    assert inside[~tail].all()
    assert (~inside[tail]).mean() > 0.9  # almost every tail column is outside somewhere
