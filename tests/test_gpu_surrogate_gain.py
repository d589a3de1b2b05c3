"""What the trained surrogate gains over the method without it.

A small instance of the whole pipeline on the GPU:
1. 2048 columns (K = 30) spun up cold for 60 days by the plain integration;
2. the network trained on those (inputs -> log10 profiles) by the
   library's own kernels;
3. 128 fresh columns (layers U{20..40}): their own 60-day cold spin-up is
   the reference; from the cold state and from the surrogate's warm start
   the plain integration runs 2 days; the distance to the reference (top
   fluxes, log10 profiles) is compared, with the integration work.
Tolerances: the warm start must be at least 3x closer to the spun-up state
than the cold start after the same 2 days, in the top fluxes (median over
columns and dissolved variables) and in the log10 profiles (median).
"""
import numpy as np
import pytest

from sedcol import capi, spinup, surrogate

pytestmark = pytest.mark.gpu

DAY = 86400.0
T_SPIN = 60.0  # days
SOLUTES = [5, 6, 7, 8, 9]


def _spun(dom, days):
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, days * DAY)
        x = g.get_state()
        _, f = g.get_rhs(x)
    return x, f


def test_warm_start_beats_cold_start(gpu_ready):
    rng = np.random.Generator(np.random.Philox(31))
    s_tr = spinup.draw_inputs(rng, 2048)
    x_tr, _ = _spun(spinup.build(s_tr, 30), T_SPIN)
    model, hist = surrogate.fit(spinup.as_matrix(s_tr), spinup.to_targets(x_tr), epochs=300, batch=256,
                                lr=2e-3, lr_final=5e-5, seed=3)
    print("val loss", hist["val_loss"][0], "->", hist["val_loss"][-1], flush=True)
    assert hist["val_loss"][-1] < 0.2 * hist["val_loss"][0]

    s = spinup.draw_inputs(rng, 128)
    K = rng.integers(20, 41, size=128)
    dom = spinup.build(s, K)
    x_ref, f_ref = _spun(dom, T_SPIN)
    x_warm0 = spinup.from_targets(surrogate.predict(model, spinup.as_matrix(s)), s["depth"], K)
    out = {}
    for name, x0 in (("cold", dom.conc0), ("warm", x_warm0)):
        with capi.SedimentColumns.from_domain(dom) as g:
            g.set_state(x0)
            g.integrate(0.0, 2.0 * DAY)
            x = g.get_state()
            _, f = g.get_rhs(x)
            d = g.diagnostics()
        sc = np.maximum(np.abs(f_ref), spinup.flux_floor(s)[None, :])
        ferr = np.abs(f - f_ref)[SOLUTES] / sc[SOLUTES]
        lg = lambda a: np.log10(np.maximum(a, 0.0) + spinup.EPS[:, None])
        perr = np.abs(lg(x) - lg(x_ref))
        out[name] = dict(flux=float(np.median(ferr)), prof=float(np.median(perr)),
                         attempts=int((d["accepted"] + d["rejected"]).sum()))
    print(out, flush=True)
    assert out["warm"]["flux"] * 3.0 <= out["cold"]["flux"]
    assert out["warm"]["prof"] * 3.0 <= out["cold"]["prof"]
    assert np.all(np.isfinite(x_warm0)) and np.all(x_warm0 >= 0.0)
