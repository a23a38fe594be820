"""ctypes binding of the C oracle (oracle/liboracle.so) -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIB = None

dp = C.POINTER(C.c_double)


class BmResult(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "qnet", "inal_auc", "ical_auc", "dvmdt_repol", "dvmdt_max", "vm_peak", "vm_valley",
        "vm_dia", "apd90", "apd50", "ca_peak", "ca_valley", "ca_dia", "cad90", "cad50")] + [
        ("steepest_pace", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Ord11CellCfg(C.Structure):
    _fields_ = [("celltype", C.c_double), ("conc", C.c_double), ("bcl", C.c_double),
                ("ic50", C.c_double * 14), ("time_point", C.c_double), ("max_dt", C.c_double),
                ("dutta", C.c_int), ("nbeats", C.c_int), ("nwin", C.c_int)]


class BdfStats(C.Structure):
    _fields_ = [(n, C.c_long) for n in ("nst", "nfe", "nsetups", "nje", "nni", "ncfn", "netf")]


class OrdCellCfg(C.Structure):
    _fields_ = [("celltype", C.c_double), ("conc", C.c_double), ("bcl", C.c_double),
                ("stim_amp_scale", C.c_double), ("dt", C.c_double), ("rtol", C.c_double),
                ("atol", C.c_double), ("herg", C.c_double * 19), ("gscale", C.c_double * 7),
                ("nbeats", C.c_int), ("nwin", C.c_int), ("maxord", C.c_int),
                ("v_apd90", C.c_double), ("v_apd50", C.c_double), ("t_first", C.c_double)]


DR_FIELDS = ("dh_apd", "dh_t_dep", "dh_t_rep", "dh_nrep", "dh_qnet", "dh_inal_auc", "dh_ical_auc",
             "bp_qnet", "bp_max_dvmdt", "bp_vm_peak", "bp_vm_dia", "bp_apd90", "bp_apd50", "bp_ca_peak",
             "bp_ca_dia", "bp_cad90", "bp_cad50")
DR_N = len(DR_FIELDS)


class DrState(C.Structure):
    _fields_ = [("bcl", C.c_double), ("dt", C.c_double), ("v_apd90", C.c_double), ("v_apd50", C.c_double),
                ("v_prev", C.c_double), ("nbeats", C.c_int), ("imax", C.c_int), ("w_prev", C.c_int),
                ("w_curr", C.c_int), ("w_cadia", C.c_int)] + [
        (n, C.c_double) for n in ("dh_t_dep", "dh_t_rep", "dh_apd", "dh_q0", "dh_n0", "dh_c0", "dh_q1",
                                  "dh_n1", "dh_c1")] + [("dh_nrep", C.c_int)] + [
        (n, C.c_double) for n in ("bp_t_dep", "bp_t_rep", "bp_apd90", "bp_apd50", "bp_q0", "bp_q1",
                                  "bp_ca_dia", "bp_ca_peak", "bp_ca_amp", "bp_t1_cad90", "bp_cad90",
                                  "bp_t1_cad50", "bp_cad50", "bp_max_dvmdt", "bp_vm_peak", "bp_vm_dia")] + [
        (n, C.c_int) for n in ("bp_ca90_inc", "bp_cad90_found", "bp_ca50_inc", "bp_cad50_found")]


class TnState(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t_last", "v_apd90", "v_prev", "t_dep", "t_rep", "apd", "sum_ca")] + [
        ("m", C.c_int)]


class TtCellCfg(C.Structure):
    _fields_ = [("bcl", C.c_double), ("dt", C.c_double), ("stim_start_unused", C.c_double),
                ("celltype", C.c_int), ("nbeats", C.c_int), ("nwin", C.c_int),
                ("conc", C.c_double), ("ic50", C.c_double * 3), ("hill", C.c_double * 3),
                ("gscale", C.c_double * 10)]


BDF_RHS = C.CFUNCTYPE(None, C.c_double, dp, dp, C.c_void_p)


class BdfOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("hmax", C.c_double),
                ("hmin", C.c_double), ("maxord", C.c_int), ("max_steps", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = C.CDLL(path)
        L.ord_init_consts.argtypes = [dp, dp, C.c_double]
        L.ord_apply_drug.argtypes = [dp, dp, C.c_double, dp]
        L.ord_rates.argtypes = [C.c_double, dp, dp, dp, dp]
        L.tt_init_consts.argtypes = [dp, dp, C.c_int]
        L.tt_rates.argtypes = [C.c_double, dp, dp, dp, dp, C.c_int]
        L.tt_rush_larsen.argtypes = [C.c_double, dp, dp, dp, dp]
        L.oracle_ord_run.argtypes = [C.POINTER(OrdCellCfg), dp, C.c_int, C.POINTER(BmResult), dp,
                                     C.POINTER(BdfStats)]
        L.oracle_tt_run.argtypes = [C.POINTER(TtCellCfg), dp, C.c_int, C.POINTER(BmResult), dp]
        L.oracle_ord_run_many.argtypes = [C.POINTER(OrdCellCfg), C.c_int, C.c_int,
                                          C.POINTER(BmResult), dp]
        L.oracle_ord_run_many.restype = C.c_double
        L.oracle_ord_run_many_from.argtypes = [C.POINTER(OrdCellCfg), C.c_int, C.c_int, dp,
                                               C.POINTER(BmResult)]
        L.oracle_ord_run_many_from.restype = C.c_double
        L.oracle_ord_run_rules.argtypes = [C.POINTER(OrdCellCfg), dp, C.c_int, C.POINTER(BmResult), dp,
                                           C.POINTER(BdfStats), dp]
        L.oracle_tt_run_rules.argtypes = [C.POINTER(TtCellCfg), dp, C.c_int, C.POINTER(BmResult), dp,
                                          C.c_double, dp]
        L.dr_init.argtypes = [C.POINTER(DrState), C.c_double, C.c_double, C.c_int, C.c_double, C.c_double]
        L.dr_begin.argtypes = [C.POINTER(DrState), C.c_double]
        L.dr_sample.argtypes = [C.POINTER(DrState), C.c_int, C.c_double, C.c_double, C.c_double, dp, C.c_double]
        L.dr_result.argtypes = [C.POINTER(DrState), dp]
        L.dr_qinward.argtypes = [C.c_double] * 4
        L.dr_qinward.restype = C.c_double
        L.tn_init.argtypes = [C.POINTER(TnState), C.c_double, C.c_int, C.c_double, C.c_double]
        L.tn_sample.argtypes = [C.POINTER(TnState), C.c_double, C.c_double, C.c_double]
        L.tn_result.argtypes = [C.POINTER(TnState), dp]
        L.oracle_tt_run_many.argtypes = [C.POINTER(TtCellCfg), C.c_int, C.c_int,
                                         C.POINTER(BmResult), dp]
        L.oracle_tt_run_many.restype = C.c_double
        L.ord11_init_consts.argtypes = [dp, dp, C.c_double]
        L.ord11_apply_dutta.argtypes = [dp]
        L.ord11_apply_drug.argtypes = [dp, C.c_double, dp]
        L.ord11_rates.argtypes = [C.c_double, dp, dp, dp, dp]
        L.ord11_rates_ex.argtypes = [C.c_double, dp, dp, dp, dp, C.c_int]
        L.ord11_rush_larsen.argtypes = [C.c_double, dp, dp, dp, dp]
        L.ord11_time_step.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
        L.ord11_time_step.restype = C.c_double
        L.oracle_ord11_run.argtypes = [C.POINTER(Ord11CellCfg), dp, C.c_int, dp]
        L.oracle_ord11_run.restype = C.c_long
        L.oracle_ord11_run_many.argtypes = [C.POINTER(Ord11CellCfg), C.c_int, C.c_int, dp, dp, dp]
        L.oracle_ord11_run_many.restype = C.c_double
        L.oracle_ord_rates_many.argtypes = [C.c_int, dp, dp, dp, C.c_double]
        L.oracle_ord_rates_many.restype = C.c_double
        L.bdf_create.argtypes = [C.c_int, BDF_RHS, C.c_void_p, C.POINTER(BdfOpts)]
        L.bdf_create.restype = C.c_void_p
        L.bdf_init.argtypes = [C.c_void_p, C.c_double, dp]
        L.bdf_advance.argtypes = [C.c_void_p, C.c_double, dp]
        L.bdf_free.argtypes = [C.c_void_p]
        L.bdf_get_stats.argtypes = [C.c_void_p, C.POINTER(BdfStats)]
        L.bdf_q.argtypes = [C.c_void_p]
        L.bdf_h.argtypes = [C.c_void_p]
        L.bdf_h.restype = C.c_double
        L.bm_create.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        L.bm_create.restype = C.c_void_p
        L.bm_begin.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.bm_sample.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 7
        L.bm_result.argtypes = [C.c_void_p, C.POINTER(BmResult)]
        L.bm_free.argtypes = [C.c_void_p]
        L.fem_q1_element.argtypes = [C.c_double] * 5 + [dp, dp]
        L.fem_q4_element.argtypes = [dp, dp] + [C.c_double] * 3 + [dp, dp]
        L.fem_hex8_element.argtypes = [dp, dp, dp, dp]
        L.fem_tet4_element.argtypes = [dp, dp, dp, dp]
        L.cg_solve_csr.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp, dp,
                                   C.c_double, C.c_int, C.c_int]
        _LIB = L
    return _LIB


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(dp)


def ord_init(celltype, conc=0.0, herg=None):
    Cc = np.zeros(206); S = np.zeros(52)
    lib().ord_init_consts(ptr(Cc), ptr(S), float(celltype))
    if herg is not None:
        h = np.ascontiguousarray(herg, dtype=np.float64)
        lib().ord_apply_drug(ptr(Cc), ptr(S), float(conc), ptr(h))
    else:
        S[43] = conc
    return Cc, S


def ord_rates(t, Cc, S):
    R = np.zeros(52); A = np.zeros(200)
    S = np.ascontiguousarray(S, dtype=np.float64)
    lib().ord_rates(float(t), ptr(Cc), ptr(R), ptr(S), ptr(A))
    return R, A


def tt_init(celltype):
    Cc = np.zeros(53); S = np.zeros(19)
    lib().tt_init_consts(ptr(Cc), ptr(S), int(celltype))
    return Cc, S


def tt_rates(t, Cc, S, celltype):
    R = np.zeros(19); A = np.zeros(70)
    S = np.ascontiguousarray(S, dtype=np.float64)
    lib().tt_rates(float(t), ptr(Cc), ptr(R), ptr(S), ptr(A), int(celltype))
    return R, A


def make_ord_cfg(celltype=0, conc=0.0, herg=None, gscale=None, bcl=1000.0, nbeats=1, nwin=None,
                 dt=0.1, rtol=1e-6, atol=1e-8, stim_amp_scale=1.0, maxord=5, v_apd90=-40.0, v_apd50=-40.0,
                 t_first=0.0):
    c = OrdCellCfg()
    c.v_apd90 = v_apd90; c.v_apd50 = v_apd50; c.t_first = t_first
    c.celltype = celltype; c.conc = conc; c.bcl = bcl; c.stim_amp_scale = stim_amp_scale
    c.dt = dt; c.rtol = rtol; c.atol = atol; c.nbeats = nbeats
    c.nwin = nbeats if nwin is None else nwin; c.maxord = maxord
    h = np.zeros(19) if herg is None else herg
    for k in range(19):
        c.herg[k] = float(h[k])
    g = np.ones(7) if gscale is None else gscale
    for k in range(7):
        c.gscale[k] = float(g[k])
    return c


def ord_run(cfg: OrdCellCfg, states=None, trace=False):
    S = np.zeros(52) if states is None else np.array(states, dtype=np.float64)
    res = BmResult(); st = BdfStats()
    nsamp = int(cfg.bcl / cfg.dt)
    vm = np.zeros(nsamp) if trace else None
    rc = lib().oracle_ord_run(C.byref(cfg), ptr(S), 0 if states is None else 1, C.byref(res),
                              ptr(vm) if trace else None, C.byref(st))
    return rc, S, res.as_dict(), vm, {n: getattr(st, n) for n, _ in st._fields_}


def ord_run_rules(cfg: OrdCellCfg, states=None, trace=False):
    """oracle_ord_run plus the driver rules: (rc, final states, biomarkers, Vm trace, rules dict)."""
    S = np.zeros(52) if states is None else np.array(states, dtype=np.float64)
    res = BmResult(); st = BdfStats()
    vm = np.zeros(int(cfg.bcl / cfg.dt)) if trace else None
    rules = np.zeros(DR_N)
    rc = lib().oracle_ord_run_rules(C.byref(cfg), ptr(S), 0 if states is None else 1, C.byref(res),
                                    ptr(vm) if trace else None, C.byref(st), ptr(rules))
    return rc, S, res.as_dict(), vm, dict(zip(DR_FIELDS, rules))


def tt_run_rules(cfg, v_apd90=-40.0):
    """oracle_tt_run plus population_tn2006's rules: (final states, biomarkers, (apd, mean Ca_i))."""
    S = np.zeros(19)
    res = BmResult()
    rules = np.zeros(2)
    lib().oracle_tt_run_rules(C.byref(cfg), ptr(S), 0, C.byref(res), None, float(v_apd90), ptr(rules))
    return S, res.as_dict(), rules


def make_tt_cfg(celltype=0, conc=0.0, ic50=(0, 0, 0), hill=(0, 0, 0), gscale=None, bcl=1000.0,
                nbeats=1, nwin=None, dt=0.02):
    c = TtCellCfg()
    c.bcl = bcl; c.dt = dt; c.celltype = int(celltype); c.nbeats = nbeats
    c.nwin = nbeats if nwin is None else nwin; c.conc = conc
    for k in range(3):
        c.ic50[k] = float(ic50[k]); c.hill[k] = float(hill[k])
    g = np.ones(10) if gscale is None else gscale
    for k in range(10):
        c.gscale[k] = float(g[k])
    return c


def tt_run(cfg: TtCellCfg, states=None, trace=False):
    S = np.zeros(19) if states is None else np.array(states, dtype=np.float64)
    res = BmResult()
    nsamp = int(round(cfg.bcl / cfg.dt))
    vm = np.zeros(nsamp) if trace else None
    lib().oracle_tt_run(C.byref(cfg), ptr(S), 0 if states is None else 1, C.byref(res),
                        ptr(vm) if trace else None)
    return S, res.as_dict(), vm


ORD11_NS, ORD11_NC, ORD11_NA, ORD11_NBM = 41, 144, 198, 16
ORD11_BM_FIELDS = ("qnet_ap", "qnet4_ap", "inal_auc_ap", "ical_auc_ap", "qnet_cl", "qnet4_cl",
                   "inal_auc_cl", "ical_auc_cl", "vm_peak", "vm_valley", "dvmdt_repol", "t_peak",
                   "vm_repol30", "vm_repol50", "vm_repol90", "steepest_pace")


def ord11_init(celltype=0.0, conc=0.0, ic50=None, dutta=False):
    Cc = np.zeros(ORD11_NC); S = np.zeros(ORD11_NS)
    lib().ord11_init_consts(ptr(Cc), ptr(S), float(celltype))
    if dutta:
        lib().ord11_apply_dutta(ptr(Cc))
    if ic50 is not None:
        a = np.ascontiguousarray(ic50, dtype=np.float64)
        lib().ord11_apply_drug(ptr(Cc), float(conc), ptr(a))
    return Cc, S


def ord11_rates(t, Cc, S, tissue_stim=-1):
    R = np.zeros(ORD11_NS); A = np.zeros(ORD11_NA)
    S = np.ascontiguousarray(S, dtype=np.float64)
    lib().ord11_rates_ex(float(t), ptr(Cc), ptr(R), ptr(S), ptr(A), int(tissue_stim))
    return R, A


def make_ord11_cfg(conc=0.0, ic50=None, bcl=1000.0, nbeats=1, nwin=1, dutta=False, celltype=0.0,
                   time_point=25.0, max_dt=1.0):
    c = Ord11CellCfg()
    c.celltype = float(celltype); c.conc = float(conc); c.bcl = float(bcl)
    if ic50 is not None:
        for k in range(14):
            c.ic50[k] = float(ic50[k])
    c.time_point = time_point; c.max_dt = max_dt
    c.dutta = int(bool(dutta)); c.nbeats = int(nbeats); c.nwin = int(nwin)
    return c


def ord11_run(cfg, states=None):
    """One cell through the DrugSimulation loop: (final states, feature record, steps)."""
    S = np.zeros(ORD11_NS) if states is None else np.ascontiguousarray(states, dtype=np.float64).copy()
    res = np.zeros(ORD11_NBM)
    n = lib().oracle_ord11_run(C.byref(cfg), ptr(S), int(states is not None), ptr(res))
    return S, res, n


def ord11_run_many(cfgs, states_in=None, nthreads=None):
    """Cells of the DrugSimulation loop over host threads: (final states, records, wall s)."""
    n = len(cfgs)
    arr = (Ord11CellCfg * n)(*cfgs)
    res = np.zeros((n, ORD11_NBM))
    sout = np.zeros((n, ORD11_NS))
    sin = None if states_in is None else np.ascontiguousarray(np.broadcast_to(states_in, (n, ORD11_NS)), dtype=np.float64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    wall = lib().oracle_ord11_run_many(arr, n, nt, None if sin is None else ptr(sin), ptr(res), ptr(sout))
    return sout, res, wall
