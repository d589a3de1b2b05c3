"""The learned warm start on the device: sedsurr_predict_device leaves the
network's output on the GPU and sedcol_set_state_profiles turns it into the
columns' state there.  Checked against the host path (sedsurr_predict,
then spinup.from_targets, the host interpolation): the state within 1e-12
relative (the two paths evaluate the geometric node
grid with different pow implementations, which may round differently in the
last bit), on ragged columns of every lane bin with land columns, which keep
their state.  Error paths are rejected with SEDCOL_EINVAL."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from sedcol import capi, spinup, surrogate

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1.0e-12, 1.0e-14
MODEL = Path(surrogate.__file__).resolve().parent / "models" / "spinup_mlp.npz"


def _case(n=3000, seed=5):
    rng = np.random.Generator(np.random.Philox(seed))
    s = spinup.draw_inputs(rng, n, tails=0.05)
    K = rng.integers(20, 41, size=n)
    K[:12] = [1, 2, 5, 12, 16, 17, 31, 32, 33, 48, 63, 64]  # every lane bin and its edges
    dom = spinup.build(s, K)
    land = np.zeros(n, dtype=np.int32)
    land[7::11] = 1
    dom.mask[:] = land
    colid = np.repeat(np.arange(n), dom.nlayers)
    dom.conc0[:, land[colid] != 0] = 1.0e20
    return s, K, dom, land[colid] != 0


@pytest.fixture(scope="module")
def model():
    return surrogate.Model.load(MODEL)


def test_device_warm_start_matches_host_path(gpu_ready, model):
    s, K, dom, land_cell = _case()
    Xn = (spinup.as_matrix(s) - model.x_mean) / model.x_std
    with surrogate.Network(model.shape, max_batch=4096) as net:
        net.set_params(model.params)
        yn_host = net.predict(Xn)
        yn_dev = net.predict_device(Xn)
        assert (yn_dev.rows, yn_dev.ld) == (3000, 330) and yn_dev.ptr != 0
        ref = spinup.from_targets(yn_host.astype(np.float64) * model.y_std + model.y_mean, s["depth"], K)
        for depth in (s["depth"], None):
            with capi.SedimentColumns.from_domain(dom) as g:
                if depth is not None:
                    surrogate.warm_init(model, yn_dev, depth)(g)
                else:
                    g.set_state_profiles(yn_dev, spinup.K_SURR, model.y_std, model.y_mean, spinup.EPS)
                x = g.get_state()
            _compare(x, ref, land_cell, RTOL if depth is not None else 1e-10, depth is not None)


def _compare(x, ref, land_cell, rtol, given):
    """Land cells untouched, wet cells within rtol of the host path."""
    wet = ~land_cell
    assert np.all(x[:, land_cell] == 1.0e20)
    err = np.abs(x[:, wet] - ref[:, wet])
    tol = rtol * np.abs(ref[:, wet]) + ATOL
    print(f"depth {'given' if given else 'from dz'}: max rel "
          f"{np.max(err / np.maximum(np.abs(ref[:, wet]), 1e-300)):.2e}", flush=True)
    assert np.all(err <= tol)
    assert np.all(np.isfinite(x[:, wet])) and np.all(x[:, wet] >= 0.0)


def test_warm_start_then_integration_runs(gpu_ready, model):
    """One simulated hour from the device warm start: finite, non-negative."""
    s, K, dom, land_cell = _case(n=500, seed=9)
    Xn = (spinup.as_matrix(s) - model.x_mean) / model.x_std
    with surrogate.Network(model.shape, max_batch=4096) as net:
        net.set_params(model.params)
        yn = net.predict_device(Xn)
        r = spinup.spin(dom, 3600.0 / spinup.YEAR, 1e-300, chunk_days=3600.0 / spinup.DAY,
                        init=surrogate.warm_init(model, yn, s["depth"]))
    x = r["state"]
    assert r["init_s"] > 0.0
    assert np.all(np.isfinite(x[:, ~land_cell])) and np.all(x[:, ~land_cell] >= 0.0)


def test_warm_start_errors(gpu_ready, model):
    s, K, dom, _ = _case(n=64, seed=2)
    Xn = (spinup.as_matrix(s) - model.x_mean) / model.x_std
    sc, of = model.y_std, model.y_mean
    host = np.zeros((64, 330), dtype=np.float32)
    L = capi.lib()
    pd = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    with surrogate.Network(model.shape, max_batch=256) as net:
        net.set_params(model.params)
        with pytest.raises(surrogate.SurrogateError):
            net.predict_device(Xn[:0])  # no rows
        yn = net.predict_device(Xn)
        with capi.SedimentColumns.from_domain(dom) as g:
            x_before = g.get_state()
            calls = [
                (host.ctypes.data, 330, 30, 1e-3, None),  # host pointer
                (yn.ptr, 330, 1, 1e-3, None),             # too few nodes
                (yn.ptr, 329, 30, 1e-3, None),            # row stride too short
                (yn.ptr, 330, 30, 0.0, None),             # dz1 not positive
                (yn.ptr, 330, 30, 1e-3, np.zeros(64)),    # depth not positive
            ]
            for prof, ld, nz, dz1, depth in calls:
                rc = L.sedcol_set_state_profiles(g._h, ctypes.c_void_p(prof), ld, nz, pd(sc), pd(of),
                                                 pd(spinup.EPS), dz1, pd(depth) if depth is not None else None)
                assert rc == capi.SEDCOL_EINVAL, (ld, nz, dz1)
            with pytest.raises(ValueError):
                g.set_state_profiles(capi.DeviceRows(yn.ptr, 63, 330), 30, sc, of, spinup.EPS)
            assert np.array_equal(g.get_state(), x_before)
