"""MI355X-native engine for the cardiac cell-model integration of the CiPA drug-population path.

The compute path is the C-ABI library ``libcepgpu.so`` (HIP kernels for gfx950, declared in
``include/cep_cellsim.h``).  This module is a thin ctypes binding of that ABI plus a
host-side mirror of the reference's CellML cell-model interface (``CellBatch``).  There is
no CPU fallback: if the library is missing the import of the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import workload  # noqa: F401  (re-exported)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libcepgpu.so")

MODEL_ORD_CIPA = 1
MODEL_TT06 = 2
MODEL_ORD2011 = 3
OPT_DUTTA = 1
OPT_TIME_POINT = 2
OPT_INTEGRATOR = 3
OPT_TIME_SERIES = 4
OPT_ORD11_WAVES = 5
OPT_DRIVER_RULES = 6
OPT_V_APD90 = 7
OPT_V_APD50 = 8
OPT_FIRST_OUTPUT = 9
OPT_KT = 10
INTEGRATOR_RUSH_LARSEN = 0
INTEGRATOR_ADAMS = 1
PREC_FP64 = 0
PREC_MIXED = 1
ST_OK = 0
ST_IDLE = 5          # a parked cell (park_cells): no job in it, runs skip it
# launch shapes (cep_set_launch_shape / cep_autotune); 0 is automatic everywhere
SHAPE_WAVES_PER_EU = 0
SHAPE_CELLS_PER_WAVE = 1
SHAPE_BLOCK = 2
SHAPE_GROUP_ORDER = 3
SHAPE_FIELDS = ("waves_per_eu", "cells_per_wave", "block", "group_order")
ORDER_STORAGE = 1
ORDER_LONGEST_FIRST = 2
STATIC_RATES_MAX_CELLS = 2048

BM_FIELDS = ("qnet", "inal_auc", "ical_auc", "dvmdt_repol", "dvmdt_max", "vm_peak", "vm_valley",
             "vm_dia", "apd90", "apd50", "ca_peak", "ca_valley", "ca_dia", "cad90", "cad50",
             "steepest_pace")
NBM = len(BM_FIELDS)
# ORd 2011 (DrugSimulation driver) record, same 16-slot width
BM11_FIELDS = ("qnet_ap", "qnet4_ap", "inal_auc_ap", "ical_auc_ap", "qnet_cl", "qnet4_cl",
               "inal_auc_cl", "ical_auc_cl", "vm_peak", "vm_valley", "dvmdt_repol", "t_peak",
               "vm_repol30", "vm_repol50", "vm_repol90", "steepest_pace")
# records of the other reference drivers' rules (cep_get_driver_rules)
DR_FIELDS = ("dh_apd", "dh_t_dep", "dh_t_rep", "dh_nrep", "dh_qnet", "dh_inal_auc", "dh_ical_auc",
             "bp_qnet", "bp_max_dvmdt", "bp_vm_peak", "bp_vm_dia", "bp_apd90", "bp_apd50", "bp_ca_peak",
             "bp_ca_dia", "bp_cad90", "bp_cad50")
DR_TT06_FIELDS = ("apd", "mean_cai")
STATUS_NAMES = {0: "ok", 1: "conv_fail", 2: "err_fail", 3: "nonfinite", 4: "max_steps"}

EXPORTED = (
    "cep_batch_create", "cep_batch_destroy", "cep_last_error", "cep_batch_info", "cep_set_cells",
    "cep_set_model_option", "cep_run_outputs",
    "cep_set_run_config", "cep_init_consts", "cep_get_constants", "cep_compute_rates",
    "cep_run_beats", "cep_run_beats_async", "cep_sync", "cep_get_states", "cep_set_states",
    "cep_get_biomarkers", "cep_get_status", "cep_get_stats", "cep_enable_vm_trace",
    "cep_get_vm_trace", "cep_get_time_series", "cep_get_dvmdt_repol_history",
    "cep_kernel_time", "cep_query", "cep_batch_device_bytes", "cep_get_driver_rules", "cep_get_qinward",
    "cep_restart_cells", "cep_park_cells",
    "cep_integrate_bdf", "cep_step_rl", "cep_compute_rates_static",
    "cep_set_launch_shape", "cep_get_launch_shape", "cep_autotune", "cep_tune_lookup",
    "cep_tissue_create", "cep_tissue_create_model", "cep_tissue_destroy", "cep_tissue_last_error", "cep_tissue_set_cells",
    "cep_tissue_set_conductivity", "cep_tissue_set_stimulus", "cep_tissue_step", "cep_tissue_diffuse",
    "cep_tissue_get_vm", "cep_tissue_get_states", "cep_tissue_set_states", "cep_tissue_time",
    "cep_tissue_kernel_times",
    "cep_mesh_read_inp", "cep_mesh_create", "cep_mesh_destroy", "cep_mesh_info", "cep_mesh_get_nodes",
    "cep_mesh_get_elements", "cep_mesh_get_celltype", "cep_mesh_get_targets", "cep_read_pacing_file",
    "cep_tissue_create_mesh", "cep_tissue_write_checkpoint", "cep_tissue_read_checkpoint",
    "cep_tissue_set_s1s2", "cep_tissue_s2_time", "cep_tissue_set_conductivity_tensor",
    "cep_release_cached_memory",
    "cep_read_ic50_csv", "cep_read_states_file", "cep_write_states_file", "cep_param_defaults",
    "cep_read_input_deck", "cep_param_concentrations", "cep_write_cipa_features", "cep_write_ord11_qni",
    "cep_write_time_series",
)
SPLIT_GODUNOV = 0
SPLIT_STRANG = 1

_lib = None
dp = C.POINTER(C.c_double)


class CepError(RuntimeError):
    pass


def library(path: str | None = None):
    """Load libcepgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise CepError(f"{p} is missing: build it with __graft_entry__.build() "
                       "(the HIP path has no CPU fallback)")
    L = C.CDLL(p)
    L.cep_batch_create.argtypes = [C.c_int, C.c_int, C.c_int]
    L.cep_batch_create.restype = C.c_void_p
    L.cep_batch_destroy.argtypes = [C.c_void_p]
    L.cep_batch_destroy.restype = None
    L.cep_release_cached_memory.argtypes = []
    L.cep_last_error.argtypes = [C.c_void_p]
    L.cep_last_error.restype = C.c_char_p
    L.cep_batch_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 5
    L.cep_set_cells.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, C.c_int, dp, C.c_int, dp]
    L.cep_set_run_config.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.c_int, C.c_double]
    L.cep_init_consts.argtypes = [C.c_void_p]
    L.cep_set_model_option.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.cep_run_outputs.argtypes = [C.c_void_p, C.c_int]
    L.cep_get_constants.argtypes = [C.c_void_p, C.c_int, dp]
    L.cep_compute_rates.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]
    L.cep_compute_rates_static.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]
    L.cep_run_beats.argtypes = [C.c_void_p, C.c_int]
    L.cep_run_beats_async.argtypes = [C.c_void_p, C.c_int]
    L.cep_sync.argtypes = [C.c_void_p]
    L.cep_get_states.argtypes = [C.c_void_p, dp]
    L.cep_set_states.argtypes = [C.c_void_p, dp]
    L.cep_get_biomarkers.argtypes = [C.c_void_p, dp]
    L.cep_get_status.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.cep_get_stats.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.cep_enable_vm_trace.argtypes = [C.c_void_p, C.c_int]
    L.cep_get_vm_trace.argtypes = [C.c_void_p, dp, C.POINTER(C.c_int)]
    L.cep_get_time_series.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]
    L.cep_get_dvmdt_repol_history.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), dp,
                                              C.POINTER(C.c_int)]
    L.cep_kernel_time.argtypes = [C.c_void_p, dp, dp]
    L.cep_get_driver_rules.argtypes = [C.c_void_p, dp]
    L.cep_integrate_bdf.argtypes = [C.c_void_p, C.c_double]
    L.cep_step_rl.argtypes = [C.c_void_p, C.c_double, C.c_double]
    L.cep_get_qinward.argtypes = [C.c_void_p, C.POINTER(C.c_int), dp]
    L.cep_query.argtypes = [C.c_void_p]
    L.cep_restart_cells.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int]
    L.cep_park_cells.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
    L.cep_set_launch_shape.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.cep_get_launch_shape.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.cep_autotune.argtypes = [C.c_void_p, C.c_int, C.c_char_p, dp, C.c_int, C.POINTER(C.c_int)]
    L.cep_tune_lookup.argtypes = [C.c_void_p, C.c_char_p]
    L.cep_batch_device_bytes.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
    L.cep_batch_device_bytes.restype = C.c_longlong
    L.cep_tissue_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
    L.cep_tissue_create.restype = C.c_void_p
    L.cep_tissue_create_model.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]
    L.cep_tissue_create_model.restype = C.c_void_p
    ip = C.POINTER(C.c_int)
    L.cep_mesh_read_inp.argtypes = [C.c_char_p]
    L.cep_mesh_read_inp.restype = C.c_void_p
    L.cep_mesh_create.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_int, ip, ip]
    L.cep_mesh_create.restype = C.c_void_p
    L.cep_mesh_destroy.argtypes = [C.c_void_p]
    L.cep_mesh_destroy.restype = None
    L.cep_mesh_info.argtypes = [C.c_void_p, ip, ip, ip, ip, ip]
    L.cep_mesh_get_nodes.argtypes = [C.c_void_p, dp]
    L.cep_mesh_get_elements.argtypes = [C.c_void_p, ip]
    L.cep_mesh_get_celltype.argtypes = [C.c_void_p, ip]
    L.cep_mesh_get_targets.argtypes = [C.c_void_p, ip]
    L.cep_read_pacing_file.argtypes = [C.c_char_p, ip, C.c_int, ip]
    L.cep_tissue_create_mesh.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.cep_tissue_create_mesh.restype = C.c_void_p
    L.cep_tissue_write_checkpoint.argtypes = [C.c_void_p, C.c_char_p]
    L.cep_tissue_set_s1s2.argtypes = [C.c_void_p, C.c_int, ip, C.c_int, C.c_int]
    L.cep_tissue_s2_time.argtypes = [C.c_void_p, dp]
    L.cep_tissue_set_conductivity_tensor.argtypes = [C.c_void_p, dp, C.c_int]
    L.cep_tissue_read_checkpoint.argtypes = [C.c_void_p, C.c_char_p]
    L.cep_tissue_destroy.argtypes = [C.c_void_p]
    L.cep_tissue_destroy.restype = None
    L.cep_tissue_last_error.argtypes = [C.c_void_p]
    L.cep_tissue_last_error.restype = C.c_char_p
    L.cep_tissue_set_cells.argtypes = [C.c_void_p, dp, dp, dp, C.c_int, dp, C.c_int]
    L.cep_tissue_set_conductivity.argtypes = [C.c_void_p, dp, dp, dp]
    L.cep_tissue_set_stimulus.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.c_double]
    L.cep_tissue_step.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                  C.POINTER(C.c_longlong)]
    L.cep_tissue_diffuse.argtypes = [C.c_void_p, C.c_double, dp, dp, C.c_double, C.c_int, C.POINTER(C.c_int)]
    L.cep_tissue_get_vm.argtypes = [C.c_void_p, dp]
    L.cep_tissue_get_states.argtypes = [C.c_void_p, dp]
    L.cep_tissue_set_states.argtypes = [C.c_void_p, dp]
    L.cep_tissue_time.argtypes = [C.c_void_p]
    L.cep_tissue_time.restype = C.c_double
    L.cep_tissue_kernel_times.argtypes = [C.c_void_p, dp, C.POINTER(C.c_longlong)]
    if path is None:
        _lib = L
    return L


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(dp)


class CellBatch:
    """Host mirror of the reference's CellML model object, for N cells at once.

    ``init_consts`` ~ ``initConsts(celltype, conc, herg, drug)`` per cell, ``compute_rates``
    ~ ``computeRates(TIME, CONSTANTS, RATES, STATES, ALGEBRAIC)``, ``run_beats`` replaces the
    per-cell CVode / solveAnalytical loops of the drivers, ``states``/``constants`` expose the
    STATES / CONSTANTS arrays in the reference layout.
    """

    def __init__(self, model: int, ncells: int, precision: int = PREC_FP64):
        self.L = library()
        h = self.L.cep_batch_create(model, ncells, precision)
        if not h:
            raise CepError(f"cep_batch_create({model}, {ncells}, {precision}) failed")
        self.h = h
        self.model = model
        self.n = ncells
        self.precision = precision
        info = [C.c_int() for _ in range(5)]
        self.L.cep_batch_info(h, *[C.byref(x) for x in info])
        _, _, self.nstates, self.nconsts, self.nalg = (x.value for x in info)

    def _chk(self, rc):
        if rc != 0:
            raise CepError(self.L.cep_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.cep_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- configuration
    def set_cells(self, celltype=None, conc=None, herg=None, gscale=None, bcl=None, first=0):
        def arr(x, w=None):
            if x is None:
                return None, 0
            a = np.ascontiguousarray(x, dtype=np.float64)
            return a, (a.shape[1] if a.ndim == 2 else 0)
        ct, _ = arr(celltype)
        cc, _ = arr(conc)
        hg, hs = arr(herg)
        gs, gss = arr(gscale)
        bc, _ = arr(bcl)
        count = next(len(a) for a in (ct, cc, hg, gs, bc) if a is not None)
        self._chk(self.L.cep_set_cells(self.h, first, count, _p(ct), _p(cc), _p(hg), hs, _p(gs), gss, _p(bc)))

    def set_run_config(self, dt=None, rtol=1e-6, atol=1e-8, maxord=5, nbeats_total=1, nwin=250,
                       stim_amp_scale=1.0):
        if dt is None:
            dt = {MODEL_ORD_CIPA: 0.1, MODEL_TT06: 0.02, MODEL_ORD2011: 1.0}[self.model]
        self._chk(self.L.cep_set_run_config(self.h, dt, rtol, atol, maxord, nbeats_total, nwin, stim_amp_scale))

    def set_model_option(self, option: int, value: float):
        """ORd 2011: OPT_DUTTA (the reference's Is_Dutta), OPT_TIME_POINT (fixed-step window),
        OPT_INTEGRATOR; ORd-CiPA: OPT_TIME_SERIES (record the selected beat's time series)."""
        self._chk(self.L.cep_set_model_option(self.h, int(option), float(value)))

    def init_consts(self):
        self._chk(self.L.cep_init_consts(self.h))

    def set_driver_rules(self, on=True, v_apd90=None, v_apd50=None, first_output=None):
        """Also evaluate the biomarker rules of the drugs_herg / benchpatch drivers (ORd-CiPA) or
        of population_tn2006 (TT06); call before init_consts."""
        self.set_model_option(OPT_DRIVER_RULES, 1.0 if on else 0.0)
        if v_apd90 is not None:
            self.set_model_option(OPT_V_APD90, v_apd90)
        if v_apd50 is not None:
            self.set_model_option(OPT_V_APD50, v_apd50)
        if first_output is not None:
            self.set_model_option(OPT_FIRST_OUTPUT, first_output)

    def driver_rules(self) -> dict:
        names = DR_TT06_FIELDS if self.model == MODEL_TT06 else DR_FIELDS
        out = np.zeros((self.n, len(names)))
        self._chk(self.L.cep_get_driver_rules(self.h, _p(out)))
        return {k: out[:, i].copy() for i, k in enumerate(names)}

    def qinward(self, control) -> np.ndarray:
        """qinward of each cell against control[i], its concentration-0 cell (-1: NaN), on the device."""
        ctl = np.ascontiguousarray(control, dtype=np.int32)
        out = np.zeros(self.n)
        self._chk(self.L.cep_get_qinward(self.h, ctl.ctypes.data_as(C.POINTER(C.c_int)), _p(out)))
        return out

    # ---- CellML arrays
    def constants(self, cell: int) -> np.ndarray:
        out = np.zeros(self.nconsts)
        self._chk(self.L.cep_get_constants(self.h, cell, _p(out)))
        return out

    def compute_rates(self, t: float, states: np.ndarray, algebraic: bool = False):
        s = np.ascontiguousarray(states, dtype=np.float64).reshape(self.n, self.nstates)
        r = np.zeros_like(s)
        a = np.zeros((self.n, self.nalg)) if algebraic else None
        self._chk(self.L.cep_compute_rates(self.h, float(t), _p(s), _p(r), _p(a)))
        return (r, a) if algebraic else r

    def compute_rates_static(self, t: float, states: np.ndarray, algebraic: bool = False):
        """compute_rates through the library's fixed device array (no device allocation; at most
        STATIC_RATES_MAX_CELLS cells)."""
        s = np.ascontiguousarray(states, dtype=np.float64).reshape(self.n, self.nstates)
        r = np.zeros_like(s)
        a = np.zeros((self.n, self.nalg)) if algebraic else None
        self._chk(self.L.cep_compute_rates_static(self.h, float(t), _p(s), _p(r), _p(a)))
        return (r, a) if algebraic else r

    def states(self) -> np.ndarray:
        out = np.zeros((self.n, self.nstates))
        self._chk(self.L.cep_get_states(self.h, _p(out)))
        return out

    def set_states(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64).reshape(self.n, self.nstates)
        self._chk(self.L.cep_set_states(self.h, _p(a)))

    # ---- integration
    def run_beats(self, nbeats: int):
        self._chk(self.L.cep_run_beats(self.h, nbeats))

    def run_outputs(self, nout: int):
        """ORd-CiPA BDF / ORd 2011 Adams: advance every cell by nout output points (blocking)."""
        self._chk(self.L.cep_run_outputs(self.h, nout))

    def integrate(self, tout: float):
        """ORd-CiPA: CVode(tout) for every cell -- integrate to the output point tout."""
        self._chk(self.L.cep_integrate_bdf(self.h, float(tout)))

    def step_rl(self, t: float, dt: float):
        """One fixed-dt step of every cell: computeRates(t) then solveAnalytical(dt)."""
        self._chk(self.L.cep_step_rl(self.h, float(t), float(dt)))

    def run_beats_async(self, nbeats: int):
        self._chk(self.L.cep_run_beats_async(self.h, nbeats))

    def query(self) -> bool:
        """True while the last run_beats_async launch is still running."""
        rc = self.L.cep_query(self.h)
        if rc < 0:
            self._chk(rc)
        return rc == 1

    def sync(self):
        self._chk(self.L.cep_sync(self.h))

    # ---- cells restarted and parked one by one (a resident batch serving jobs)
    @staticmethod
    def _cells(idx):
        a = np.ascontiguousarray(idx, dtype=np.int32)
        return a, a.ctypes.data_as(C.POINTER(C.c_int))

    def restart_cells(self, idx, nbeats_total: int, nwin: int | None = None):
        """Restart the listed cells from their model's initial states with the parameters last
        given by set_cells, a fresh integrator and biomarker machine, and a biomarker window of
        the last nwin (default all) of nbeats_total beats from now; other cells are untouched."""
        a, p = self._cells(idx)
        nw = nbeats_total if nwin is None or nwin <= 0 else nwin
        self._chk(self.L.cep_restart_cells(self.h, p, len(a), int(nbeats_total), int(nw)))

    def park_cells(self, idx):
        """Status ST_IDLE for the listed cells: runs skip them until they are restarted."""
        a, p = self._cells(idx)
        self._chk(self.L.cep_park_cells(self.h, p, len(a)))

    def enable_vm_trace(self, on=True):
        self._chk(self.L.cep_enable_vm_trace(self.h, 1 if on else 0))

    def vm_trace(self):
        ns = C.c_int()
        self._chk(self.L.cep_get_vm_trace(self.h, None, C.byref(ns)))
        out = np.zeros((self.n, ns.value))
        self._chk(self.L.cep_get_vm_trace(self.h, _p(out), C.byref(ns)))
        return out

    def time_series(self, cell: int) -> dict:
        """The selected (steepest) beat's samples of one cell: t, vm, cai (n), ires (n x 6:
        INaL ICaL Ito IKr IKs IK1), and the selected pace (OPT_TIME_SERIES must be on)."""
        n, pace = C.c_int(), C.c_int()
        self._chk(self.L.cep_get_time_series(self.h, int(cell), 0, None, None, None, None, C.byref(n),
                                             C.byref(pace)))
        k = n.value
        t, vm, ca, ires = np.zeros(k), np.zeros(k), np.zeros(k), np.zeros((k, 6))
        self._chk(self.L.cep_get_time_series(self.h, int(cell), k, _p(t), _p(vm), _p(ca), _p(ires), C.byref(n),
                                             C.byref(pace)))
        return {"t": t, "vm": vm, "cai": ca, "ires": ires, "pace": pace.value}

    def dvmdt_repol_history(self, cell: int):
        """(paces, dvmdt_repol) of every eligible pace of the window (OPT_TIME_SERIES on)."""
        n = C.c_int()
        self._chk(self.L.cep_get_dvmdt_repol_history(self.h, int(cell), 0, None, None, C.byref(n)))
        paces = np.zeros(n.value, dtype=np.int32)
        vals = np.zeros(n.value)
        self._chk(self.L.cep_get_dvmdt_repol_history(self.h, int(cell), n.value,
                                                     paces.ctypes.data_as(C.POINTER(C.c_int)), _p(vals),
                                                     C.byref(n)))
        return paces, vals

    def biomarkers(self) -> dict:
        out = np.zeros((self.n, NBM))
        self._chk(self.L.cep_get_biomarkers(self.h, _p(out)))
        names = BM11_FIELDS if self.model == MODEL_ORD2011 else BM_FIELDS
        return {k: out[:, i].copy() for i, k in enumerate(names)}

    def biomarker_array(self) -> np.ndarray:
        out = np.zeros((self.n, NBM))
        self._chk(self.L.cep_get_biomarkers(self.h, _p(out)))
        return out

    def status(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.int32)
        self._chk(self.L.cep_get_status(self.h, out.ctypes.data_as(C.POINTER(C.c_int))))
        return out

    def stats(self) -> np.ndarray:
        out = np.zeros((self.n, 8), dtype=np.int64)
        self._chk(self.L.cep_get_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong))))
        return out

    def kernel_time(self):
        avg, tot = C.c_double(), C.c_double()
        n = self.L.cep_kernel_time(self.h, C.byref(avg), C.byref(tot))
        return n, avg.value, tot.value

    def set_launch_shape(self, knob: int, value: int):
        self._chk(self.L.cep_set_launch_shape(self.h, knob, int(value)))

    def launch_shape(self) -> dict:
        """The shape the next launch uses, automatic choices resolved."""
        out = {}
        for k, name in enumerate(SHAPE_FIELDS):
            v = C.c_int()
            self._chk(self.L.cep_get_launch_shape(self.h, k, C.byref(v)))
            out[name] = v.value
        return out

    def autotune(self, trial_beats: int = 1, cache: str | None = None):
        """Times every candidate launch shape on trial_beats beats of this batch (its device state
        restored afterwards) and keeps the fastest: (candidates with their ms, chosen index)."""
        rep = np.zeros((16, len(SHAPE_FIELDS) + 1))
        nc = C.c_int()
        best = self.L.cep_autotune(self.h, int(trial_beats), cache.encode() if cache else None, _p(rep), 16,
                                   C.byref(nc))
        self._chk(min(best, 0))
        rows = [dict(zip(SHAPE_FIELDS + ("ms",), [int(x) for x in r[:-1]] + [float(r[-1])]))
                for r in rep[:nc.value]]
        return rows, best

    def tune_lookup(self, cache: str) -> bool:
        """Applies a cached choice to the shapes still automatic; True if the cache had one."""
        rc = self.L.cep_tune_lookup(self.h, cache.encode())
        self._chk(min(rc, 0))
        return rc == 1


class Mesh:
    """A tissue mesh of the reference's file format (Cmesh::loadInp, mesh.cpp:59-183), read
    from a file or given as arrays (xyz: nnodes x 3, conn: nelem x nodes_per_elem, 0-based)."""

    def __init__(self, xyz=None, conn=None, celltype=None, dimension: int = 2, path: str | None = None):
        self.L = library()
        if path is not None:
            h = self.L.cep_mesh_read_inp(path.encode())
            if not h:
                raise CepError(f"cep_mesh_read_inp({path}) failed")
        else:
            x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
            c = np.ascontiguousarray(conn, dtype=np.int32)
            ct = None if celltype is None else np.ascontiguousarray(celltype, dtype=np.int32)
            h = self.L.cep_mesh_create(len(x), _p(x), c.shape[0], c.shape[1], dimension,
                                       c.ctypes.data_as(C.POINTER(C.c_int)),
                                       None if ct is None else ct.ctypes.data_as(C.POINTER(C.c_int)))
            if not h:
                raise CepError("cep_mesh_create failed")
        self.h = h
        v = [C.c_int() for _ in range(5)]
        self.L.cep_mesh_info(h, *[C.byref(q) for q in v])
        self.nnodes, self.nelem, self.nodes_per_elem, self.dimension, has_ct = (q.value for q in v)
        self.has_celltype = bool(has_ct)

    @classmethod
    def read_inp(cls, path: str) -> "Mesh":
        return cls(path=path)

    def nodes(self) -> np.ndarray:
        out = np.zeros((self.nnodes, 3))
        self.L.cep_mesh_get_nodes(self.h, _p(out))
        return out

    def elements(self) -> np.ndarray:
        out = np.zeros((self.nelem, self.nodes_per_elem), dtype=np.int32)
        self.L.cep_mesh_get_elements(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def celltype(self):
        if not self.has_celltype:
            return None
        out = np.zeros(self.nnodes, dtype=np.int32)
        self.L.cep_mesh_get_celltype(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def targets(self) -> np.ndarray:
        out = np.zeros(3, dtype=np.int32)
        self.L.cep_mesh_get_targets(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.cep_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Tissue:
    """Monodomain tissue: the reference's Cheart reaction step + Cfem diffusion step
    (heart.cpp, fem.cpp) on the device, on a structured Q1 grid (nx x ny elements of hx x hy
    mm) or on a mesh of S4 quadrilaterals, C3D8I hexahedra or C3D4 tetrahedra
    (Tissue.from_mesh)."""

    def __init__(self, nx: int, ny: int, hx: float, hy: float, precision: int = PREC_FP64,
                 model: int = MODEL_ORD_CIPA, _handle=None, _nn=None, _ne=None):
        self.L = library()
        h = _handle if _handle is not None else self.L.cep_tissue_create_model(nx, ny, hx, hy, model, precision)
        if not h:
            raise CepError(f"cep_tissue_create({nx}, {ny}, {hx}, {hy}) failed")
        self.h = h
        self.nx, self.ny = nx, ny
        self.nn = _nn if _nn is not None else (nx + 1) * (ny + 1)
        self.ne = _ne if _ne is not None else nx * ny
        self.model = model
        self.dimension = 2
        self.nstates = 41 if model == MODEL_ORD2011 else 52

    @classmethod
    def from_mesh(cls, mesh: "Mesh", precision: int = PREC_FP64, model: int = MODEL_ORD_CIPA) -> "Tissue":
        L = library()
        h = L.cep_tissue_create_mesh(mesh.h, model, precision)
        if not h:
            raise CepError("cep_tissue_create_mesh failed (S4, C3D8I or C3D4 elements with a positive "
                           "Jacobian at every Gauss point are required)")
        t = cls(0, 0, 0.0, 0.0, precision, model, _handle=h, _nn=mesh.nnodes, _ne=mesh.nelem)
        t.dimension = mesh.dimension
        return t

    def set_conductivity_tensor(self, d):
        """Per-element tensor: shape (ne,) or a scalar = isotropic DE (fem.cpp:162-170);
        (ne, 3) = 2-D (xx, xy, yy); (ne, 6) = 3-D (xx, xy, xz, yy, yz, zz)."""
        a = np.asarray(d, dtype=np.float64)
        if a.ndim == 0:
            a = np.full(self.ne, float(a))
        ncomp = 1 if a.ndim == 1 else a.shape[1]
        a = np.ascontiguousarray(a.reshape(self.ne, ncomp))
        self._chk(self.L.cep_tissue_set_conductivity_tensor(self.h, _p(a), ncomp))

    def set_s1s2(self, npace1: int, s2_nodes=(), target: int = 0):
        """S1 pacing stops after npace1 paces; during the last one the S2 nodes are reset to
        their initial states once the target node has repolarised (heart.cpp:315-323,475-501)."""
        idx = np.ascontiguousarray(np.asarray(s2_nodes, dtype=np.int32))
        self._chk(self.L.cep_tissue_set_s1s2(self.h, int(npace1), idx.ctypes.data_as(C.POINTER(C.c_int)), len(idx),
                                             int(target)))

    def s2_time(self) -> float:
        v = C.c_double()
        self._chk(self.L.cep_tissue_s2_time(self.h, C.byref(v)))
        return v.value

    def write_checkpoint(self, path: str):
        self._chk(self.L.cep_tissue_write_checkpoint(self.h, path.encode()))

    def read_checkpoint(self, path: str):
        self._chk(self.L.cep_tissue_read_checkpoint(self.h, path.encode()))

    def _chk(self, rc):
        if rc != 0:
            raise CepError(self.L.cep_tissue_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.cep_tissue_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def node(self, ix: int, iy: int) -> int:
        return iy * (self.nx + 1) + ix

    def set_cells(self, celltype=0.0, conc=0.0, herg=None, gscale=None):
        n = self.nn
        ct = np.ascontiguousarray(np.broadcast_to(np.asarray(celltype, dtype=np.float64), (n,)))
        cc = np.ascontiguousarray(np.broadcast_to(np.asarray(conc, dtype=np.float64), (n,)))
        w = 14 if self.model == MODEL_ORD2011 else 19
        hg = None if herg is None else np.ascontiguousarray(np.broadcast_to(np.asarray(herg, np.float64), (n, w)))
        gs = None if gscale is None else np.ascontiguousarray(np.broadcast_to(np.asarray(gscale, np.float64), (n, 7)))
        self._chk(self.L.cep_tissue_set_cells(self.h, _p(ct), _p(cc), _p(hg), w, _p(gs), 7))

    def set_conductivity(self, sxx, sxy, syy):
        ne = self.ne
        a = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (ne,))) for v in (sxx, sxy, syy)]
        self._chk(self.L.cep_tissue_set_conductivity(self.h, _p(a[0]), _p(a[1]), _p(a[2])))

    def set_stimulus(self, nodes, amp_scale=1.0, bcl=1000.0, start=0.0, duration=0.5):
        idx = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32))
        self._chk(self.L.cep_tissue_set_stimulus(self.h, idx.ctypes.data_as(C.POINTER(C.c_int)), len(idx),
                                                 amp_scale, bcl, start, duration))

    def step(self, nsteps: int, dt: float, split: int = SPLIT_GODUNOV, cg_rtol: float = 1e-5, cg_maxit: int = 200):
        it = C.c_longlong()
        self._chk(self.L.cep_tissue_step(self.h, nsteps, dt, split, cg_rtol, cg_maxit, C.byref(it)))
        return it.value

    def diffuse(self, dt: float, v, cg_rtol: float = 1e-5, cg_maxit: int = 500):
        a = np.ascontiguousarray(v, dtype=np.float64)
        x = np.empty(self.nn)
        it = C.c_int()
        self._chk(self.L.cep_tissue_diffuse(self.h, dt, _p(a), _p(x), cg_rtol, cg_maxit, C.byref(it)))
        return x, it.value

    def vm(self) -> np.ndarray:
        out = np.empty(self.nn)
        self._chk(self.L.cep_tissue_get_vm(self.h, _p(out)))
        return out

    def states(self) -> np.ndarray:
        out = np.empty((self.nn, self.nstates))
        self._chk(self.L.cep_tissue_get_states(self.h, _p(out)))
        return out

    def set_states(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64)
        self._chk(self.L.cep_tissue_set_states(self.h, _p(a)))

    @property
    def time(self) -> float:
        return self.L.cep_tissue_time(self.h)

    def kernel_times(self):
        """(avg ms, launches) of the reaction, SpMV and CG-update kernels since the last call."""
        avg = np.zeros(3)
        n = (C.c_longlong * 3)()
        self._chk(self.L.cep_tissue_kernel_times(self.h, _p(avg), n))
        return avg, np.array(list(n))


def release_cached_memory() -> None:
    """Return the device memory kept from destroyed batches (it is otherwise reused by later
    batches); waits for the device."""
    if library().cep_release_cached_memory() != 0:
        raise CepError("cep_release_cached_memory failed")


def device_bytes(model: int, ncells: int, max_bcl: float, dt: float, vm_trace: bool = False) -> int:
    """Device memory a batch of `ncells` cells takes (memory-based admission)."""
    v = library().cep_batch_device_bytes(model, ncells, float(max_bcl), float(dt), int(vm_trace))
    if v < 0:
        raise CepError("cep_batch_device_bytes: bad arguments")
    return int(v)


def qinward(inal_auc_drug, ical_auc_drug, inal_auc_ctl, ical_auc_ctl):
    """qInward = (INaL_AUC_drug/INaL_AUC_control + ICaL_AUC_drug/ICaL_AUC_control) / 2
    (CEPCell_NEO/modules/cipa_drug_bench.cpp:372-377)."""
    return 0.5 * (np.asarray(inal_auc_drug) / inal_auc_ctl + np.asarray(ical_auc_drug) / ical_auc_ctl)
