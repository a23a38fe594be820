"""cep_compute_rates_static (the computeRates entry point that allocates nothing) against
cep_compute_rates, through the C ABI: bitwise equal rates and ALGEBRAIC rows for ORd-CiPA, TT06
and ORd 2011 in both precisions, on two calls in a row on the same GPU, the second with a smaller
batch (the fixed device array keeps the first call's contents, so stale rows would show); a batch
above the array's size is refused with a clear error."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
cep = importlib.import_module("2ddrugsimulation_amd")


def _batch(model, n, prec, rng):
    b = cep.CellBatch(model, n, prec)
    if model == cep.MODEL_ORD_CIPA:
        wl = cep.workload.ord_population(20251204, 2, 2, (n + 3) // 4)
        b.set_cells(celltype=wl["celltype"][:n], conc=wl["conc"][:n], herg=wl["herg"][:n],
                    gscale=wl["gscale"][:n], bcl=wl["bcl"][:n])
        b.set_run_config(dt=0.1)
    elif model == cep.MODEL_TT06:
        pop = cep.workload.tt_population(20251205, n)
        b.set_cells(celltype=pop["celltype"], conc=pop["conc"], herg=np.concatenate([pop["ic50"], pop["hill"]], 1),
                    gscale=pop["gscale"], bcl=pop["bcl"])
        b.set_run_config(dt=0.02)
    else:
        ic50 = np.array([5000, 1.0, 3000, 1.0, 20000, 1.0, 800, 1.0, 10000, 1.0, 4000, 1.0, 900, 0.8])
        b.set_cells(celltype=np.zeros(n), conc=rng.uniform(0, 1000, n), herg=np.repeat(ic50[None], n, 0),
                    bcl=np.full(n, 1000.0))
        b.set_run_config(dt=1.0)
    b.init_consts()
    s = b.states()
    s[:, 0] = rng.uniform(-90.0, 40.0, n)         # a spread of voltages, every current path
    return b, s


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("model", [1, 2, 3])
def test_static_rates_equal_the_allocating_entry_point(model, prec):
    rng = np.random.default_rng(10 * model + prec)
    for n in (300, 37):                           # the second call smaller than the first
        b, s = _batch(model, n, prec, rng)
        t = 3.1
        r_ref, a_ref = b.compute_rates(t, s, algebraic=True)
        r, a = b.compute_rates_static(t, s, algebraic=True)
        r_only = b.compute_rates_static(t, s)
        b.close()
        np.testing.assert_array_equal(r, r_ref)
        np.testing.assert_array_equal(a, a_ref)
        np.testing.assert_array_equal(r_only, r_ref)
        assert np.isfinite(r).any()


def test_static_rates_refuse_a_batch_beyond_the_array():
    n = cep.STATIC_RATES_MAX_CELLS + 1
    with cep.CellBatch(cep.MODEL_TT06, n, cep.PREC_FP64) as b:
        b.set_cells(celltype=np.zeros(n), conc=np.zeros(n), herg=np.zeros((n, 6)), gscale=np.ones((n, 10)),
                    bcl=np.full(n, 1000.0))
        b.set_run_config(dt=0.02)
        b.init_consts()
        s = b.states()
        with pytest.raises(cep.CepError, match="at most 2048 cells"):
            b.compute_rates_static(0.0, s)
        r = b.compute_rates(0.0, s)           # the allocating entry point still takes it
        assert np.isfinite(r).all()
