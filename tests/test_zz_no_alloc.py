"""sedcol_rhs_noalloc (the allocation-free get_rhs) against sedcol_rhs,
through the C ABI: bitwise equal right-hand sides and top fluxes on
consecutive calls that share the device's fixed workspace (the second on a
smaller domain, then the first domain again), in the packed and the Fortran
layout, land columns included; a domain beyond the workspace is refused
with SEDCOL_EINVAL before anything runs."""
import numpy as np
import pytest

from sedcol import cases, capi

pytestmark = pytest.mark.gpu


def _state(dom, seed):
    rng = np.random.default_rng(seed)
    x = dom.conc0.copy()
    wet = np.repeat(dom.mask == 0, dom.nlayers)
    x[0:5, wet] = rng.uniform(0.0, 1.0e4, size=(5, int(wet.sum())))
    x[8, wet] = rng.uniform(0.0, 50.0, size=int(wet.sum()))
    x[:, wet] *= rng.uniform(0.5, 1.5, size=(11, int(wet.sum())))
    return x


def _pair(g, x):
    r1, f1 = g.get_rhs(x)
    r2, f2 = g.get_rhs_noalloc(x)
    return r1, f1, r2, f2


def test_noalloc_matches_rhs_on_consecutive_calls(gpu_ready):
    big = cases.shelf_sea(n=24, days=1, seed=4)   # 576 columns, land included
    small = cases.plain_columns(ni=3, nj=1, K=17, seed=2)
    assert big.mask.any() and (big.mask == 0).any()
    xb, xs = _state(big, 1), _state(small, 2)
    with capi.SedimentColumns.from_domain(big) as gb, capi.SedimentColumns.from_domain(small) as gs:
        for g, x, name in ((gb, xb, "first"), (gs, xs, "second, smaller"), (gb, xb, "first again")):
            r1, f1, r2, f2 = _pair(g, x)
            print(f"{name}: {g.ncol} columns, rhs equal {np.array_equal(r1, r2)}, fluxes equal "
                  f"{np.array_equal(f1, f2)}", flush=True)
            assert np.array_equal(r1, r2)
            assert np.array_equal(f1, f2)
        land = np.repeat(big.mask != 0, big.nlayers)
        r2, _ = gb.get_rhs_noalloc(xb)
        assert np.all(r2[:, land] == 0.0)


def test_noalloc_fortran_layout(gpu_ready):
    dom = cases.plain_columns(ni=3, nj=2, K=20, seed=6)
    K, ncol = 20, dom.ncol
    cellf = lambda a: np.ascontiguousarray(a.reshape(ncol, K).T)
    x = dom.fortran_state(_state(dom, 3))
    with capi.SedimentColumns(np.full(ncol, K), layout="fortran") as g:
        g.set_params(dom.params.si())
        g.set_geometry(cellf(dom.dz), cellf(dom.dzc), cellf(dom.phi), cellf(dom.phiI), cellf(dom.bfac), dom.biot)
        g.set_forcing(*dom.record(0))
        g.set_state(x)
        r1, f1, r2, f2 = _pair(g, x)
    assert np.array_equal(r1, r2)
    assert np.array_equal(f1, f2)


def test_noalloc_refuses_a_domain_beyond_its_workspace(gpu_ready):
    dom = cases.plain_columns(ni=150, nj=100, K=30, seed=0)  # 450 000 cells, 79 MB of staging
    with capi.SedimentColumns.from_domain(dom) as g:
        with pytest.raises(capi.SedcolError) as e:
            g.get_rhs_noalloc(dom.conc0)
        assert e.value.code == capi.SEDCOL_EINVAL
        assert "workspace" in str(e.value)
        r1, _ = g.get_rhs(dom.conc0)  # the allocating entry point still serves it
        assert np.all(np.isfinite(r1))
