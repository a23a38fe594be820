"""Pin the oracle's biomarker state machine (CEPCell_NEO/modules/cipa_drug_bench.cpp:210-381)
on synthetic traces whose answers follow from their geometry.

Each beat of the trace is a piecewise-linear action potential (rest -88 mV, upstroke from the
stimulus time to +40 mV in 2 ms, linear repolarisation at a per-beat slope back to rest) and a
triangular Ca transient.  The expected values are derived from that geometry: APD90/APD50 from
where the linear repolarisation crosses the lower edge of the reference's 2 mV band, the peak
sampled 6 ms after the stimulus, CaD50/CaD90 from the triangle's crossing times, and the winning
beat from the largest dV/dt in the repolarisation range (the least negative slope).
"""
from __future__ import annotations

import ctypes
import math

import pytest

import oracle_lib as O

BCL, DT, STIM = 100.0, 0.1, 3.0
V_REST, V_TOP, UP_MS = -88.0, 40.0, 2.0
CA0, CAP, CA_RISE_END, CA_FALL_END = 1.0e-4, 1.0e-3, 13.05, 63.05


def beat_trace(tb, slope):
    """Vm, dVm/dt and Ca at beat-relative time tb for a repolarisation slope (mV/ms)."""
    if tb < STIM:
        v, dv = V_REST, 0.0
    elif tb < STIM + UP_MS:
        dv = (V_TOP - V_REST) / UP_MS
        v = V_REST + dv * (tb - STIM)
    else:
        v = V_TOP - slope * (tb - STIM - UP_MS)
        dv = -slope
        if v <= V_REST:
            v, dv = V_REST, 0.0
    if tb < STIM:
        ca = CA0
    elif tb < CA_RISE_END:
        ca = CA0 + (CAP - CA0) * (tb - STIM) / (CA_RISE_END - STIM)
    elif tb < CA_FALL_END:
        ca = CAP - (CAP - CA0) * (tb - CA_RISE_END) / (CA_FALL_END - CA_RISE_END)
    else:
        ca = CA0
    return v, dv, ca


def run_machine(slopes, nwin, qsum=1.7):
    nb = len(slopes)
    steps = int(round(BCL / DT))
    L = O.lib()
    h = L.bm_create(BCL, DT, STIM, nb, nwin, 1)
    L.bm_begin(h, V_REST, CA0)
    for iout in range(1, nb * steps + 1):
        t = iout * DT
        p = iout // steps
        p = min(p, nb - 1)
        v, dv, ca = beat_trace(t - BCL * p, slopes[p])
        L.bm_sample(h, iout, t, v, ca, dv, qsum, 0.0, 0.0)
    r = O.BmResult()
    L.bm_result(h, ctypes.byref(r))
    L.bm_free(h)
    return r


def expected_for_slope(s):
    """Closed-form answers for one beat with repolarisation slope s."""
    t_peak = STIM + 6.0                                   # Vm sampled 6 ms after the stimulus
    v_peak = V_TOP - s * (t_peak - STIM - UP_MS)
    amp = v_peak - V_REST                                 # valley = Vm at the beat start (rest)
    out = {"vm_peak": v_peak, "dvmdt_repol": -s, "dvmdt_max": (V_TOP - V_REST) / UP_MS}
    for name, frac in (("apd90", 0.9), ("apd50", 0.5)):
        level = v_peak - frac * amp
        # the last sample with level-2 < V < level lies within one dt before V reaches level-2
        t_lower = STIM + UP_MS + (V_TOP - (level - 2.0)) / s
        out[name] = t_lower - STIM
    # Ca: 50% / 90% levels of the sampled peak; durations between the crossings
    # the sampled peak is the larger of the two samples either side of the triangle's apex
    t_a = math.floor(CA_RISE_END / DT) * DT
    t_b = t_a + DT
    ca_peak = max(CA0 + (CAP - CA0) * (t_a - STIM) / (CA_RISE_END - STIM),
                  CAP - (CAP - CA0) * (t_b - CA_RISE_END) / (CA_FALL_END - CA_RISE_END))
    for name, frac in (("cad50", 0.5), ("cad90", 0.9)):
        level = ca_peak - frac * (ca_peak - CA0)
        t_up = STIM + (level - CA0) / (CAP - CA0) * (CA_RISE_END - STIM)
        t_dn = CA_RISE_END + (CAP - level) / (CAP - CA0) * (CA_FALL_END - CA_RISE_END)
        out[name] = t_dn - t_up
    out["ca_peak"] = ca_peak
    return out


def test_apd_and_peak_of_a_single_window_beat():
    s = 3.17
    r = run_machine([3.3, s], nwin=1)
    e = expected_for_slope(s)
    assert r.steepest_pace == 1
    assert r.vm_peak == pytest.approx(e["vm_peak"], abs=1e-9)
    assert r.vm_valley == V_REST
    assert r.dvmdt_max == pytest.approx(e["dvmdt_max"])
    assert r.dvmdt_repol == pytest.approx(e["dvmdt_repol"])
    for k in ("apd90", "apd50"):
        # the sample lies in the last dt before the band's lower edge is crossed
        assert e[k] - DT - 1e-9 <= getattr(r, k) < e[k] + 1e-9, (k, getattr(r, k), e[k])
    assert r.vm_dia == V_REST and r.ca_dia == CA0


def test_steepest_beat_selection_picks_largest_repolarisation_dvdt():
    slopes = [3.3, 3.17, 2.03, 4.11]          # window = beats 1..3; dV/dt -3.17, -2.03, -4.11
    r = run_machine(slopes, nwin=3)
    assert r.steepest_pace == 2
    e = expected_for_slope(2.03)
    assert r.dvmdt_repol == pytest.approx(-2.03)
    assert e["apd90"] - DT - 1e-9 <= r.apd90 < e["apd90"] + 1e-9
    assert e["apd50"] - DT - 1e-9 <= r.apd50 < e["apd50"] + 1e-9
    # a larger slope gives a shorter APD: the pick is not the last or the first window beat
    e_last = expected_for_slope(4.11)
    assert r.apd90 > e_last["apd90"] + 1.0


def test_calcium_transient_durations_and_peak():
    r = run_machine([3.3, 3.17, 2.03], nwin=2)
    e = expected_for_slope(2.03)
    assert r.ca_peak == pytest.approx(e["ca_peak"], rel=1e-12)
    assert r.ca_valley == CA0
    for k in ("cad50", "cad90"):
        # both crossing times are resolved to one sample each
        assert abs(getattr(r, k) - e[k]) <= 2 * DT + 1e-9, (k, getattr(r, k), e[k])


def test_qnet_integrates_one_beat():
    r = run_machine([3.3, 3.17, 2.03], nwin=2, qsum=1.7)
    # rectangle rule over the BCL/dt samples of one beat
    assert r.qnet == pytest.approx(1.7 * BCL, rel=1e-12)


def test_window_without_eligible_ap_keeps_first_window_beat():
    """A beat whose sampled peak is below 0 mV never becomes eligible: APD stays 0 and the
    first window beat is reported (the reference's initialisation of the result)."""
    slopes = [3.3, 12.0, 12.5]                # peak 6 ms after the stimulus is 40-12*4 < 0
    r = run_machine(slopes, nwin=2)
    assert r.steepest_pace == 1
    assert r.apd90 == 0.0 and r.apd50 == 0.0
    assert r.vm_peak < 0.0
