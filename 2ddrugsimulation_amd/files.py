"""The reference drivers' input and output files, through the library's C ABI
(include/cep_io.h): IC50 sample CSV, EDISON input deck, steady-state dumps, and the CiPA and
ORd 2011 result files."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import NBM, CepError, _p, library


class Param(C.Structure):
    """cep_param_t: the driver parameters of the EDISON deck."""
    _fields_ = [("simulation_mode", C.c_int), ("celltype", C.c_int), ("is_dutta", C.c_int),
                ("is_print_graph", C.c_int), ("is_using_output", C.c_int), ("pace_max", C.c_int),
                ("last_drug_check_pace", C.c_int), ("bcl", C.c_double), ("dt", C.c_double),
                ("dt_write", C.c_double), ("inet_vm_threshold", C.c_double),
                ("drug_name", C.c_char * 100), ("concs", C.c_char * 100)]

    def concentrations(self) -> np.ndarray:
        out = np.zeros(64)
        n = C.c_int()
        _chk(_lib().cep_param_concentrations(C.byref(self), _p(out), len(out), C.byref(n)), "concentrations")
        return out[:n.value].copy()


def _lib():
    L = library()
    if not getattr(L, "_io_typed", False):
        ip = C.POINTER(C.c_int)
        L.cep_read_ic50_csv.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int, ip]
        L.cep_read_states_file.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int, ip]
        L.cep_write_states_file.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_int]
        L.cep_param_defaults.argtypes = [C.POINTER(Param)]
        L.cep_param_defaults.restype = None
        L.cep_read_input_deck.argtypes = [C.c_char_p, C.POINTER(Param)]
        L.cep_param_concentrations.argtypes = [C.POINTER(Param), C.POINTER(C.c_double), C.c_int, ip]
        L.cep_write_cipa_features.argtypes = [C.c_char_p, ip, C.POINTER(C.c_double), ip, C.c_int,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.cep_write_ord11_qni.argtypes = [C.c_char_p, ip, C.POINTER(C.c_double), C.c_int,
                                          C.POINTER(C.c_double), C.c_int]
        L.cep_write_time_series.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, ip,
                                            C.POINTER(C.c_double), C.c_int, C.c_int]
        L._io_typed = True
    return L


def _chk(rc, what):
    if rc != 0:
        raise CepError(f"{what} failed ({rc})")


def _ints(a):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=np.int32)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_int))


def read_ic50_csv(path: str, max_rows: int = 100000) -> np.ndarray:
    rows = np.zeros((max_rows, 14))
    n = C.c_int()
    _chk(_lib().cep_read_ic50_csv(path.encode(), _p(rows), max_rows, C.byref(n)), f"reading {path}")
    return rows[:n.value].copy()


def read_states(path: str, max_n: int = 256) -> np.ndarray:
    out = np.zeros(max_n)
    n = C.c_int()
    _chk(_lib().cep_read_states_file(path.encode(), _p(out), max_n, C.byref(n)), f"reading {path}")
    return out[:n.value].copy()


def write_states(path: str, states) -> None:
    s = np.ascontiguousarray(states, dtype=np.float64).ravel()
    _chk(_lib().cep_write_states_file(path.encode(), _p(s), len(s)), f"writing {path}")


def read_input_deck(path: str | None = None) -> Param:
    p = Param()
    _lib().cep_param_defaults(C.byref(p))
    if path is not None:
        _chk(_lib().cep_read_input_deck(path.encode(), C.byref(p)), f"reading {path}")
    return p


def write_cipa_features(prefix: str, sample_ids, bm, status=None, inal_ctl=None, ical_ctl=None,
                        header: bool = True) -> None:
    ids, idp = _ints(sample_ids)
    b = np.ascontiguousarray(bm, dtype=np.float64).reshape(-1, NBM)
    st, stp = _ints(status)
    ia = None if inal_ctl is None else np.ascontiguousarray(inal_ctl, dtype=np.float64)
    ic = None if ical_ctl is None else np.ascontiguousarray(ical_ctl, dtype=np.float64)
    _chk(_lib().cep_write_cipa_features(prefix.encode(), idp, _p(b), stp, len(ids), _p(ia), _p(ic), int(header)),
         "writing CiPA features")


def write_ord11_qni(prefix: str, sample_ids, bm11, ctl=None, header: bool = True) -> None:
    ids, idp = _ints(sample_ids)
    b = np.ascontiguousarray(bm11, dtype=np.float64).reshape(-1, NBM)
    c = None if ctl is None else np.ascontiguousarray(ctl, dtype=np.float64).reshape(-1, NBM)
    _chk(_lib().cep_write_ord11_qni(prefix.encode(), idp, _p(b), len(ids), _p(c), int(header)),
         "writing ORd 2011 qNet file")


def write_time_series(prefix: str, sample_id: int, ts: dict, history=None, header: bool = True) -> None:
    """The reference's Is_Print_Graph files of one sample: <prefix>_vmcheck_smp<id>.plt,
    _ca_i_smp, _ires_smp, _max_dvdt_repol_smp.  ts: CellBatch.time_series(cell); history:
    CellBatch.dvmdt_repol_history(cell) or None."""
    t = np.ascontiguousarray(ts["t"], dtype=np.float64)
    vm = np.ascontiguousarray(ts["vm"], dtype=np.float64)
    ca = np.ascontiguousarray(ts["cai"], dtype=np.float64)
    ires = np.ascontiguousarray(ts["ires"], dtype=np.float64).reshape(-1, 6)
    if history is None:
        rp, rpp, rv, nrep = None, None, None, 0
    else:
        rp, rpp = _ints(history[0])
        rv = np.ascontiguousarray(history[1], dtype=np.float64)
        nrep = len(rp)
    _chk(_lib().cep_write_time_series(prefix.encode(), int(sample_id), _p(t), _p(vm), _p(ca), _p(ires), len(t),
                                      rpp, _p(rv), nrep, int(header)),
         "writing time series")


def read_pacing_file(path: str, max_n: int = 10000000) -> np.ndarray:
    """The S1/S2 pacing-site file (heart.cpp:966-998): 0-based node numbers."""
    L = _lib()
    L.cep_read_pacing_file.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    n = C.c_int()
    _chk(L.cep_read_pacing_file(path.encode(), None, 0, C.byref(n)), "reading pacing file")
    out = np.zeros(n.value, dtype=np.int32)
    _chk(L.cep_read_pacing_file(path.encode(), out.ctypes.data_as(C.POINTER(C.c_int)), min(n.value, max_n),
                                C.byref(n)), "reading pacing file")
    return out
