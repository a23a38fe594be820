"""The launch-shape autotuner (cep_autotune, cep_set/get_launch_shape, cep_tune_lookup): every
candidate shape computes the same results, tuning leaves the batch's state as it was, the choice
is the fastest measured one and is persisted and found again by key.  No reference counterpart:
the reference runs one cell per CPU thread and has no launch shapes."""
import importlib
import os

import numpy as np
import pytest

cep = importlib.import_module("2ddrugsimulation_amd")
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _snapshot(b):
    return b.states(), b.biomarker_array(), b.status(), b.stats()


def _assert_same(a, c):
    for x, y in zip(a, c):
        np.testing.assert_array_equal(x, y)


def _tt_batch(pop, herg, prec):
    n = len(pop["bcl"])
    b = cep.CellBatch(cep.MODEL_TT06, n, prec)
    b.set_cells(celltype=pop["celltype"], conc=pop["conc"], herg=herg, gscale=pop["gscale"], bcl=pop["bcl"])
    b.set_run_config(dt=0.02, nbeats_total=3, nwin=3)
    b.init_consts()
    return b


def _ord_batch(pop, prec):
    n = len(pop["bcl"])
    b = cep.CellBatch(cep.MODEL_ORD_CIPA, n, prec)
    b.set_cells(celltype=pop["celltype"], conc=pop["conc"], herg=pop["herg"], gscale=pop["gscale"],
                bcl=pop["bcl"])
    b.set_run_config(dt=0.1, rtol=1e-6, atol=1e-8, nbeats_total=3, nwin=3)
    b.init_consts()
    return b


@pytest.mark.gpu
def test_tt06_shapes_agree_and_tuning_is_invisible(tmp_path):
    """TT06, 3000 cells: the register split (1 or 2 waves per SIMD) and 128- or 256-thread
    workgroups give bitwise one beat of the default shape; autotune restores the
    state, picks the fastest row of its report, writes one cache line per key (a second run
    replaces it) and a new batch of the same key finds it, by call and through CEP_TUNE_CACHE."""
    pop = cep.workload.tt_population(20251210, 3000)
    herg = np.concatenate([pop["ic50"], pop["hill"]], axis=1)
    prec = cep.PREC_MIXED
    with _tt_batch(pop, herg, prec) as b:
        # automatic: 47 waves, fewer than the SIMDs, so all 512 registers; one wave per workgroup
        assert b.launch_shape() == {"waves_per_eu": 1, "cells_per_wave": 64, "block": 64, "group_order": 0}
        b.run_beats(1)
        ref = _snapshot(b)
    assert (ref[2] == 0).all()
    for wpe in (1, 2):
        for block in (128, 256):
            with _tt_batch(pop, herg, prec) as b:
                b.set_launch_shape(cep.SHAPE_WAVES_PER_EU, wpe)
                b.set_launch_shape(cep.SHAPE_BLOCK, block)
                b.run_beats(1)
                _assert_same(_snapshot(b), ref)
    cache = str(tmp_path / "tune.txt")
    with _tt_batch(pop, herg, prec) as b:
        before = _snapshot(b)
        rows, best = b.autotune(1, cache)
        _assert_same(_snapshot(b), before)
        assert len(rows) == 4 and all(r["ms"] > 0 for r in rows)
        assert best == int(np.argmin([r["ms"] for r in rows]))
        chosen = b.launch_shape()
        assert (chosen["waves_per_eu"], chosen["block"]) == (rows[best]["waves_per_eu"], rows[best]["block"])
        b.run_beats(1)
        _assert_same(_snapshot(b), ref)
        print("TT06 3000 cells:", rows, "chosen", best)
        b.autotune(1, cache)
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and "model=2 prec=1" in lines[0] and "cells<=4096" in lines[0]
    want = {k: v for k, v in rows[best].items() if k != "ms"}
    with _tt_batch(pop, herg, prec) as b:
        assert b.tune_lookup(cache)
        got = b.launch_shape()
    with _tt_batch(pop, herg, cep.PREC_FP64) as b:
        assert not b.tune_lookup(cache)                   # another key
    os.environ["CEP_TUNE_CACHE"] = cache
    try:
        with _tt_batch(pop, herg, prec) as b:
            got_env = b.launch_shape()
    finally:
        del os.environ["CEP_TUNE_CACHE"]
    lines = open(cache).read().splitlines()
    shape = {k: v for k, v in zip(cep.SHAPE_FIELDS, [int(x.split("=")[1]) for x in lines[0].split(" | ")[1].split()])}
    assert got == shape and got_env == shape
    assert shape["cells_per_wave"] == want["cells_per_wave"] == 64


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [cep.PREC_MIXED, cep.PREC_FP64])
def test_ord_cipa_group_orders_are_bitwise_equal(prec):
    """ORd-CiPA BDF, 96 cells in 6 groups of BCL 500 / 1000 / 2000 ms stored shortest first: one beat
    with the groups dispatched in storage order and longest first gives bitwise the same states,
    biomarkers, statistics and status; autotune over the two orders leaves the state as it was."""
    pop = cep.workload.ord_population(20251211, 4, 4, 6)
    order = np.argsort(pop["bcl"], kind="stable")
    pop = {k: v[order] for k, v in pop.items()}
    out = []
    for o in (cep.ORDER_STORAGE, cep.ORDER_LONGEST_FIRST):
        with _ord_batch(pop, prec) as b:
            assert b.launch_shape()["group_order"] == cep.ORDER_STORAGE     # 6 groups all fit at once
            b.set_launch_shape(cep.SHAPE_GROUP_ORDER, o)
            assert b.launch_shape()["group_order"] == o
            b.run_beats(1)
            out.append(_snapshot(b))
    assert (out[0][2] == 0).all() and out[0][3][:, 7].min() >= 5000
    _assert_same(out[0], out[1])
    if prec == cep.PREC_MIXED:
        with _ord_batch(pop, prec) as b:
            b.run_beats(1)
            before = _snapshot(b)
            rows, best = b.autotune(1)
            _assert_same(_snapshot(b), before)
            assert [r["group_order"] for r in rows] == [cep.ORDER_STORAGE, cep.ORDER_LONGEST_FIRST]
            assert b.launch_shape()["group_order"] == rows[best]["group_order"]
            print("ORd-CiPA 96 cells:", rows)


@pytest.mark.gpu
def test_ord2011_shapes_and_tuning():
    """ORd 2011 Rush-Larsen, 200 quinidine cells from the steady state: the one-wave kernel with all
    512 registers gives bitwise the two-wave-per-SIMD build; autotune times the four shapes, restores
    the state, and its choice runs within the wavefront-split tolerance of the default (states
    1e-9, step counts 0.5 %)."""
    q = np.loadtxt(os.path.join(G, "quinidine_IC50_samples10.csv"), delimiter=",", skiprows=1)
    s0 = np.loadtxt(os.path.join(G, "ord2011_output_orudy.dat"))
    n = 200
    ic50 = np.vstack([q[i % len(q)] for i in range(n)])
    conc = np.array([(3237.0, 6474.0, 9711.0)[i % 3] for i in range(n)])

    def batch():
        b = cep.CellBatch(cep.MODEL_ORD2011, n, cep.PREC_FP64)
        b.set_cells(celltype=np.full(n, 2.0), conc=conc, herg=ic50, bcl=np.full(n, 2000.0))
        b.set_run_config(dt=1.0, nbeats_total=2, nwin=2)
        b.init_consts()
        b.set_states(np.repeat(s0[None, :], n, axis=0))
        return b

    out = {}
    for wpe in (2, 1):
        with batch() as b:
            b.set_launch_shape(cep.SHAPE_CELLS_PER_WAVE, 64)
            b.set_launch_shape(cep.SHAPE_WAVES_PER_EU, wpe)
            assert b.launch_shape()["waves_per_eu"] == wpe
            b.run_beats(1)
            out[wpe] = _snapshot(b)
    _assert_same(out[1], out[2])
    with batch() as b:
        with pytest.raises(cep.CepError):
            b.set_launch_shape(cep.SHAPE_CELLS_PER_WAVE, 8)
        with pytest.raises(cep.CepError):
            b.set_launch_shape(cep.SHAPE_GROUP_ORDER, cep.ORDER_LONGEST_FIRST)
        before = _snapshot(b)
        rows, best = b.autotune(1)
        _assert_same(_snapshot(b), before)
        assert [r["cells_per_wave"] for r in rows] == [64, 64, 32, 16]
        b.run_beats(1)
        tuned = _snapshot(b)
        print("ORd 2011 200 cells:", rows, "chosen", best)
    s1, _, st1, n1 = out[2]
    s2, _, st2, n2 = tuned
    assert (st1 == 0).all() and (st2 == 0).all()
    assert np.all(np.abs(n2[:, 0] - n1[:, 0]) <= 0.005 * n1[:, 0])
    np.testing.assert_allclose(s2, s1, rtol=1e-9, atol=1e-9)
