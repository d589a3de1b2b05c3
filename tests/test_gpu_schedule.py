"""Adaptive-EMP schedules and kernel shapes give identical results.

A column's operations do not depend on how its attempts are split over
launches (windows, captured graphs, one launch per step), on the lanes per
column (the layer reductions pad with identity values) or on the workgroup
size, so every schedule and shape is compared bitwise with the default one
(one persistent launch per layer bin), which the parity tests compare
bitwise with the oracle.  Ragged columns from all three layer bins, a
forcing series whose record boundaries fall inside windows.
"""
import dataclasses

import numpy as np
import pytest

from sedcol import cases, capi
import oracle as O

pytestmark = pytest.mark.gpu

SPAN = 9 * 3600.0


def _ragged_series(ncol=300, seed=21, nrec=4, dt_rec=3 * 3600.0):
    dom = cases.ragged_columns(ncol=ncol, seed=seed)
    bins = np.where(dom.nlayers <= 16, 16, np.where(dom.nlayers <= 32, 32, 64))
    assert set(np.unique(bins)) == {16, 32, 64}
    f = 1.0 + 0.1 * np.arange(nrec)
    temp = dom.temp[0][None, :] + 2.0 * np.arange(nrec)[:, None]
    cbw = dom.cbw[0][None] * f[:, None, None]
    fup = dom.fup[0][None] * f[:, None, None]
    return dataclasses.replace(dom, temp=temp, cbw=cbw, fup=fup, t0=0.0, dt_rec=dt_rec)


def _run(dom, schedule=None, tuning=None, days=SPAN):
    with capi.SedimentColumns.from_domain(dom) as g:
        if schedule:
            g.set_schedule(**schedule)
        if tuning:
            g.set_tuning(*tuning)
        g.integrate(0.0, days)
        out = dict(x=g.get_state(), f=g.fluxes(), b=g.budget(), d=g.diagnostics(), ms=g.last_timing())
    return out


def _assert_same(a, b):
    assert np.array_equal(a["x"], b["x"])
    assert np.array_equal(a["f"], b["f"])
    assert np.array_equal(a["b"], b["b"])
    for k in ("accepted", "rejected", "forced", "zero_sink", "newton", "dt_next", "dt_minacc", "p_min"):
        assert np.array_equal(a["d"][k], b["d"][k]), k


@pytest.fixture(scope="module")
def ragged():
    dom = _ragged_series()
    return dom, _run(dom)


def test_default_matches_oracle(gpu_ready, ragged):
    dom, ref = ragged
    orc = O.Oracle(dom, dom.params.si())
    xo = dom.conc0.copy()
    spans = cases.record_spans(0.0, SPAN, dom.nrec, dom.t0, dom.dt_rec)
    assert orc.integrate_series(xo, 0.0, SPAN, spans, nthreads=8) == 0
    assert np.array_equal(ref["x"], xo)
    assert np.array_equal(ref["d"]["accepted"], orc.diag["accepted"])
    assert np.all(ref["x"] >= 0.0)


@pytest.mark.parametrize("schedule", [
    dict(window=100, graph_windows=0, check_every=1),
    dict(window=100, graph_windows=4, check_every=1),
    dict(window=1, graph_windows=0, check_every=16),
    dict(window=7, graph_windows=3, check_every=1, concurrent_bins=1),
    dict(window=0, concurrent_bins=1),
], ids=["window100", "graph100x4", "one-launch-per-step", "graph7x3-concurrent", "concurrent"])
def test_schedule_bitwise(gpu_ready, ragged, schedule):
    dom, ref = ragged
    out = _run(dom, schedule=schedule)
    print(schedule, "ms", out["ms"], "accepted", int(out["d"]["accepted"].sum()), flush=True)
    _assert_same(out, ref)


@pytest.mark.parametrize("tuning", [
    ([16, 32, 64], [4, 2, 1]),
    ([32, 64, 64], [8, 4, 4]),
    ([64, 64, 64], [1, 2, 4]),
    ([16, 32, 64], [16, 8, 4]),
    ([16, 16, 32], [4, 4, 2], [1, 2, 2]),
    ([16, 32, 32], [16, 8, 8], [2, 2, 2]),
    ([16, 16, 64], [4, 8, 2], [1, 2, 2]),
], ids=["lanes16-32-64-wg64", "lanes32-64-64", "lanes64-all", "default", "two-layers-wg64", "two-layers-all",
        "two-layers-wide"])
def test_kernel_shape_bitwise(gpu_ready, ragged, tuning):
    """Lanes per column, workgroup size and layers per lane (two layers per
    lane: k_emp_adaptive2) all give the default shape's results."""
    dom, ref = ragged
    out = _run(dom, tuning=tuning)
    print(tuning, "ms", out["ms"], flush=True)
    _assert_same(out, ref)


def test_two_layers_fast_form_shelf_bitwise(gpu_ready):
    """The bench's arithmetic form (fast) with two layers per lane on 16
    lanes (K = 30): shelf-sea columns over two daily forcing records, equal
    to the default shape and to the oracle bit for bit."""
    full = cases.shelf_sea(n=40, days=2, seed=3)
    dom = full.subset(np.flatnonzero(full.mask == 0)[:200])
    dom.params.arith = 1
    span = 2 * 86400.0
    ref = _run(dom, days=span)
    out = _run(dom, tuning=([16, 16, 32], [16, 4, 8], [1, 2, 2]), days=span)
    _assert_same(out, ref)
    orc = O.Oracle(dom, dom.params.si())
    xo = dom.conc0.copy()
    spans = cases.record_spans(0.0, span, dom.nrec, dom.t0, dom.dt_rec)
    assert orc.integrate_series(xo, 0.0, span, spans, nthreads=8) == 0
    assert np.array_equal(out["x"], xo)
    assert np.array_equal(out["d"]["accepted"], orc.diag["accepted"])


@pytest.mark.parametrize("arith", [0, 1], ids=["reference", "fast"])
def test_two_layers_layer_count_edges(gpu_ready, arith):
    """Two layers per lane at the edges of the lane mapping: K = 1 (top and
    bottom in sub-layer 0 of lane 0), odd K (the last lane's second layer
    empty), K = 2 LPC (every lane full), in every bin and with the widest
    lanes; bitwise equal to one layer per lane."""
    from sedcol import spinup

    rng = np.random.Generator(np.random.Philox(9))
    Ks = np.array([1, 2, 3, 7, 15, 16, 17, 18, 31, 32, 33, 34, 47, 63, 64, 64])
    dom = spinup.build(spinup.draw_inputs(rng, Ks.size), Ks)
    dom.params.arith = arith
    span = 6 * 3600.0
    ref = _run(dom, days=span)
    for tuning in (([16, 16, 32], [4, 8, 4], [2, 2, 2]), ([64, 64, 64], [4, 2, 1], [2, 2, 2])):
        out = _run(dom, tuning=tuning, days=span)
        print(tuning, "accepted", int(out["d"]["accepted"].sum()), flush=True)
        _assert_same(out, ref)


@pytest.mark.parametrize("change", ["geometric", "ros2"])
def test_two_layers_fall_back_where_unsupported(gpu_ready, change):
    """A two-layers-per-lane shape installed on a handle whose parameters
    need the one-layer kernels (geometric-mean weights, the linearly
    implicit stepper) runs those, with the same results as the default
    shape."""
    dom = _ragged_series(ncol=60, seed=5)
    if change == "geometric":
        dom.params.emp_weights = 1
    else:
        dom.params.adaptive_method = 7
    span = 2 * 3600.0
    ref = _run(dom, days=span)
    out = _run(dom, tuning=([16, 16, 32], [4, 4, 2], [1, 2, 2]), days=span)
    _assert_same(out, ref)


def test_graph_replayed_across_calls(gpu_ready, ragged):
    """The captured graph is reused by later calls (new span, reordered
    queues) and must still match the direct schedule."""
    dom, _ = ragged
    res = []
    for sched in (dict(window=0), dict(window=50, graph_windows=2)):
        with capi.SedimentColumns.from_domain(dom) as g:
            g.set_schedule(**sched)
            for k in range(3):
                g.integrate(k * 3 * 3600.0, (k + 1) * 3 * 3600.0)
            res.append((g.get_state(), g.diagnostics()))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1]["accepted"], res[1][1]["accepted"])


def test_autotune_restores_state_and_installs_shape(gpu_ready, ragged):
    dom, ref = ragged
    with capi.SedimentColumns.from_domain(dom) as g:
        x0 = g.get_state()
        d0 = g.diagnostics()
        best, table = g.autotune(0.0, 1800.0, repeats=1)
        assert np.array_equal(g.get_state(), x0)
        assert np.array_equal(g.diagnostics(), d0)
        assert g.get_tuning() == best
        for b, lmin in enumerate((16, 32, 64)):
            assert best["layers_per_lane"][b] in (1, 2)
            assert best["lanes_per_column"][b] >= max(16, lmin // best["layers_per_lane"][b])
            assert best["lanes_per_column"][b] * best["columns_per_group"][b] in (64, 128, 256)
            for ip in range(2):  # both layers-per-lane choices are timed
                valid = table[b, ip][table[b, ip] >= 0.0]
                assert valid.size >= 3 and np.all(valid > 0.0)
        print("autotune", best, "\n", np.round(table, 3), flush=True)
        g.integrate(0.0, SPAN)
        out = dict(x=g.get_state(), f=g.fluxes(), b=g.budget(), d=g.diagnostics())
    _assert_same(out, ref)


def test_schedule_and_tuning_errors(gpu_ready):
    dom = cases.ragged_columns(ncol=40, seed=2)
    with capi.SedimentColumns.from_domain(dom) as g:
        for bad in (dict(window=-1), dict(check_every=0), dict(graph_windows=-2), dict(concurrent_bins=2)):
            with pytest.raises(capi.SedcolError) as e:
                g.set_schedule(**bad)
            assert e.value.code == capi.SEDCOL_EINVAL
        for lanes, cpg in (([8, 32, 64], [8, 8, 4]), ([16, 16, 64], [4, 4, 1]), ([16, 32, 64], [3, 2, 1])):
            with pytest.raises(capi.SedcolError) as e:
                g.set_tuning(lanes, cpg)
            assert e.value.code == capi.SEDCOL_EINVAL
        # layers per lane: 1 or 2, and lanes for half the layers at two
        for lanes, cpg, lpl in (([16, 32, 64], [4, 2, 1], [1, 3, 1]), ([16, 16, 16], [4, 4, 4], [1, 2, 2])):
            with pytest.raises(capi.SedcolError) as e:
                g.set_tuning(lanes, cpg, lpl)
            assert e.value.code == capi.SEDCOL_EINVAL
        g.set_tuning([16, 16, 32], [4, 4, 2], [1, 2, 2])
        assert g.get_tuning() == {"lanes_per_column": [16, 16, 32], "columns_per_group": [4, 4, 2],
                                  "layers_per_lane": [1, 2, 2]}
        assert g.get_schedule() == dict(window=0, graph_windows=0, check_every=1, concurrent_bins=0)


def test_checkpoint_continues_bitwise(gpu_ready, ragged):
    dom, ref = ragged
    half = 4 * 3600.0 + 17.0
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, half)
        blob = g.save_checkpoint()
    with capi.SedimentColumns.from_domain(dom) as g:
        g.load_checkpoint(blob)
        assert g.save_checkpoint() == blob
        g.integrate(half, SPAN)
        x = g.get_state()
        d = g.diagnostics()
    # same columns, same per-column operation sequence, split at `half`
    # instead of at a record boundary: states agree with an uninterrupted
    # run split the same way
    with capi.SedimentColumns.from_domain(dom) as g:
        g.integrate(0.0, half)
        g.integrate(half, SPAN)
        assert np.array_equal(g.get_state(), x)
        assert np.array_equal(g.diagnostics()["accepted"], d["accepted"])
    other = cases.ragged_columns(ncol=40, seed=2)
    with capi.SedimentColumns.from_domain(other) as g:
        with pytest.raises(capi.SedcolError) as e:
            g.load_checkpoint(blob)
        assert e.value.code == capi.SEDCOL_EINVAL
        with pytest.raises(capi.SedcolError):
            g.load_checkpoint(b"not a checkpoint" * 4)
