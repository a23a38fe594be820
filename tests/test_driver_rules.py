"""Pins of the oracle's driver rules (oracle/driver_rules.c): the biomarker rules of
CEP7Cell/modules/drugs_herg.cpp:221-315, CEPPatch_Population_Drugs/benchpatch.cpp:320-432 and
CEP7Cell/modules/population_tn2006.cpp:143-191, on synthetic traces whose crossing times are
known in closed form and fall strictly between output points (so the expected sample is the first
grid point after the analytical crossing)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import oracle_lib as O

BCL, DT, NB = 100.0, 0.5, 4
SLOPE_UP, SLOPE_DN = 54.0, 27.0          # mV/ms: -88 -> +20 in 2 ms, +20 -> -88 in 4 ms
CA_DECAY = 61.3                          # ms from the Ca peak back to the base


def plateau(b):
    return 40.0 + 2.0 * b                # AP plateau length of beat b (ms)


def vm(t):
    """-88 mV rest; upstroke at 10 ms into each beat, plateau at +20, linear repolarisation."""
    b = int(t // BCL)
    tau = t - b * BCL
    a = plateau(b)
    if tau < 10.0:
        return -88.0
    if tau < 12.0:
        return -88.0 + SLOPE_UP * (tau - 10.0)
    if tau < 12.0 + a:
        return 20.0
    if tau < 12.0 + a + 4.0:
        return 20.0 - SLOPE_DN * (tau - 12.0 - a)
    return -88.0


def grid_after(t):
    """first output time strictly after t (t never on the grid here)"""
    return math.ceil(t / DT) * DT


def ca_of(t):
    """triangular Ca transient: 0.1 base, peak 1 + 0.1 b at 20 ms, back to base at 81.3 ms"""
    b = int(t // BCL)
    tau = t - b * BCL
    pk = 1.0 + 0.1 * b
    if 10.0 <= tau < 20.0:
        return 0.1 + (pk - 0.1) * (tau - 10.0) / 10.0
    if 20.0 <= tau < 20.0 + CA_DECAY:
        return pk - (pk - 0.1) * (tau - 20.0) / CA_DECAY
    return 0.1


def run_rules(v_apd90=-40.0, v_apd50=0.0):
    s = O.DrState()
    L = O.lib()
    L.dr_init(C.byref(s), BCL, DT, NB, v_apd90, v_apd50)
    L.dr_begin(C.byref(s), -88.0)
    imax = int(NB * BCL / DT)
    for k in range(1, imax + 1):
        t = k * DT
        acc = np.array([float(k * k), 2.0 * k, 3.0 * k])
        dvdt = 50.0 - (t - 330.0) ** 2
        L.dr_sample(C.byref(s), k, t, vm(t), ca_of(t), O.ptr(acc), dvdt)
    out = np.zeros(O.DR_N)
    L.dr_result(C.byref(s), O.ptr(out))
    return dict(zip(O.DR_FIELDS, out))


def test_drugs_herg_apd_at_fixed_threshold():
    r = run_rules()
    b = NB - 1
    t_up = grid_after(b * BCL + 10.0 + 48.0 / SLOPE_UP)                 # -88 -> -40
    t_dn = grid_after(b * BCL + 12.0 + plateau(b) + 60.0 / SLOPE_DN)     # +20 -> -40
    assert r["dh_t_dep"] == t_up and r["dh_t_rep"] == t_dn
    assert r["dh_apd"] == t_dn - t_up
    assert r["dh_nrep"] == NB


def test_qnet_windows_of_the_two_drivers():
    """drugs_herg reads the accumulators after outputs (BCL/dt)(n-2)+BCL and (BCL/dt)(n-1)+BCL;
    benchpatch tests its counter before the increment, so it reads one output later."""
    r = run_rules()
    bs = int(BCL / DT)
    w0, w1 = bs * (NB - 2) + int(BCL), bs * (NB - 1) + int(BCL)
    assert r["dh_qnet"] == (w1 * w1 - w0 * w0) / 1000.0
    assert r["dh_inal_auc"] == 2.0 * (w1 - w0)
    assert r["dh_ical_auc"] == 3.0 * (w1 - w0)
    assert r["bp_qnet"] == ((w1 + 1) ** 2 - (w0 + 1) ** 2) / 1000.0


def test_benchpatch_apd_shares_its_depolarisation_time():
    """In the last cycle the v_apd50 upward crossing (later than the v_apd90 one) overwrites the
    depolarisation time that the APD90 uses (benchpatch.cpp:336-338, 414-416)."""
    r = run_rules(v_apd90=-40.0, v_apd50=0.0)
    b = NB - 1
    t_up50 = grid_after(b * BCL + 10.0 + 88.0 / SLOPE_UP)               # -88 -> 0
    t_dn50 = grid_after(b * BCL + 12.0 + plateau(b) + 20.0 / SLOPE_DN)   # +20 -> 0
    t_dn90 = grid_after(b * BCL + 12.0 + plateau(b) + 60.0 / SLOPE_DN)
    assert r["bp_apd50"] == t_dn50 - t_up50
    assert r["bp_apd90"] == t_dn90 - t_up50
    assert r["bp_vm_peak"] == 20.0 and r["bp_vm_dia"] == -88.0
    assert r["bp_max_dvmdt"] == 50.0                                    # at t = 330 in the last cycle


def test_benchpatch_calcium_from_the_previous_beat_amplitude():
    r = run_rules()
    bs = int(BCL / DT)
    t_dia = (bs * (NB - 2) + 1) * DT                                    # counter == (BCL/dt)(n-2)
    ca_dia = ca_of(t_dia)
    pk = 1.0 + 0.1 * (NB - 2)                                           # previous beat's peak
    assert r["bp_ca_dia"] == ca_dia and abs(r["bp_ca_peak"] - pk) < 1e-15
    amp = pk - ca_dia
    pk3 = 1.0 + 0.1 * (NB - 1)
    t0 = (NB - 1) * BCL
    for frac, key in ((0.1, "bp_cad90"), (0.5, "bp_cad50")):
        lvl = frac * amp + ca_dia
        t_rise = t0 + 10.0 + 10.0 * (lvl - 0.1) / (pk3 - 0.1)           # first sample at or above
        t_fall = t0 + 20.0 + CA_DECAY * (pk3 - lvl) / (pk3 - 0.1)       # first sample below
        assert min(t_rise % DT, t_fall % DT) > 1e-6                      # between grid points
        assert abs(r[key] - (grid_after(t_fall) - grid_after(t_rise))) < 1e-12


def test_drugs_herg_skips_the_last_output():
    """The output that ends the run breaks out of the loop before the checks (drugs_herg.cpp:226):
    a repolarisation exactly there is not counted."""
    s = O.DrState()
    L = O.lib()
    L.dr_init(C.byref(s), 10.0, 1.0, 1, -40.0, -40.0)
    L.dr_begin(C.byref(s), -88.0)
    acc = np.zeros(3)
    for k, v in enumerate([-88.0, 10.0, 10.0, 10.0, 10.0, 10.0, 10.0, 10.0, 10.0, -80.0], start=1):
        L.dr_sample(C.byref(s), k, float(k), v, 0.1, O.ptr(acc), 0.0)
    out = np.zeros(O.DR_N)
    L.dr_result(C.byref(s), O.ptr(out))
    assert out[O.DR_FIELDS.index("dh_t_dep")] == 2.0
    assert out[O.DR_FIELDS.index("dh_nrep")] == 0
    assert out[O.DR_FIELDS.index("bp_apd90")] == 8.0                   # benchpatch checks every output


def test_population_tn2006_mean_calcium_divides_by_count_minus_one():
    s = O.TnState()
    L = O.lib()
    L.tn_init(C.byref(s), BCL, 2, -40.0, -88.0)
    tcurr, n_last = 0.0, 0
    while tcurr < 2 * BCL - 1e-9:
        tcurr += DT
        last = tcurr >= BCL
        n_last += last
        L.tn_sample(C.byref(s), tcurr, vm(tcurr), 0.3 if last else 0.1)
    out = np.zeros(2)
    L.tn_result(C.byref(s), O.ptr(out))
    assert abs(out[1] - 0.3 * n_last / (n_last - 1)) < 1e-14
    t_up = grid_after(BCL + 10.0 + 48.0 / SLOPE_UP)
    t_dn = grid_after(BCL + 12.0 + plateau(1) + 60.0 / SLOPE_DN)
    assert abs(out[0] - (t_dn - t_up)) < 1e-9


def test_qinward_closed_form():
    assert O.lib().dr_qinward(2.0, 3.0, 4.0, 6.0) == 0.5


def test_ord_run_rules_leaves_the_cipa_record_unchanged():
    """Evaluating the rules does not change the integration or the CiPA biomarkers."""
    import importlib
    wl = importlib.import_module("2ddrugsimulation_amd").workload.ord_single_drug(seed=20251203)
    cfg = O.make_ord_cfg(conc=wl["conc"][1], herg=wl["herg"][1], nbeats=2, nwin=2, bcl=500.0)
    rc, S, res, _, _ = O.ord_run(cfg)
    rc2, S2, res2, _, rules = O.ord_run_rules(cfg)
    assert rc == rc2 == 0 and np.array_equal(S, S2) and res == res2
    assert rules["dh_nrep"] == 2 and 200 < rules["dh_apd"] < 300
