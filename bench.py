"""Headline benchmark: ORd-CiPA cell-beats/s on one MI355X (batched population), plus p50/p99
job latency of the serving loop.

Workload: 10,000 ORd-CiPA endo cells = 50 seeded drugs x 4 concentrations x 50 variants with
7 lognormal(0, 0.2) conductance scales, BCL drawn from {500, 1000, 2000} ms, CVODE-style BDF
(rtol 1e-6, atol 1e-8, max step = output dt = 0.1 ms), on-device biomarkers at every output
point.  A "step" is one beat of the whole population (every cell advances one beat of its
own BCL); inputs and states are resident in HBM.  The timed solver is the mixed-precision
one (FP32 gating/currents, FP64 state/Newton/LU); an FP64 run of a subsample, as one plain
job, checks APD90 and qNet and gives the plain per-job FP64 rate.  A serving session
(Poisson arrivals of mixed ORd-CiPA / TT06 jobs of 1-400 cells, 10-40 beats, BCL 500/1000/2000
ms, continuously batched into resident batches at the cell level) gives the p50/p99 latency.

`--workload ord2011` times the reference's own ORd 2011 run (quinidine, 30 cells x 1000
beats), whose wall time the reference published, and a batched population of that model.
`--workload serving` sweeps the Poisson arrival rate of the serving session (default 2, 8 and
32 jobs/s, the same seeded jobs at every rate, 10-100 beats each) and reports p50/p99 latency and
served cell-beats/s per rate; the value is the highest served rate in cell-beats/s.
`--workload tt06` times the ten Tusscher 2006 population: 200,000 cells (endo/epi/M, Hill
block of IKr/ICaL/INa, ten lognormal(0, 0.25) conductance scales), BCL 1000 ms, Rush-Larsen
dt 0.02 ms, on-device APD90/CaTD; a step is one beat of every cell.
`--workload tissue` instead times the 2-D monodomain step (512 x 512 Q1 elements, ORd-CiPA
nodes, Strang splitting, Jacobi-PCG diffusion) in node-steps/s; a "step" is one time step of
the whole tissue.

Prints one JSON line.  --gpus N > 1 is not supported (single-GPU build).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

# The serving loop keeps one HIP stream per resident batch (one per model and BCL: up to 6 here);
# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (default 4, some of them taken by the
# runtime itself), and a persistent beat kernel on a shared queue holds back the kernels queued
# behind it (measured: 4 batches' beats took twice as long as 2 with 4 queues; 8 ran side by
# side with 16 queues).
# Read by the HIP runtime when it initialises, so set before any HIP call.
SERVE_HW_QUEUES = 24
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < SERVE_HW_QUEUES:
    os.environ["GPU_MAX_HW_QUEUES"] = str(SERVE_HW_QUEUES)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
cep = importlib.import_module("2ddrugsimulation_amd")
serving = importlib.import_module("2ddrugsimulation_amd.serving")

# Algorithmic work per unit of the BDF kernel (flops: + - * / each 1; transcendentals apart).
# RHS: counted by oracle/opcount.cpp on the ORd-CiPA model (1292 flops, 172 exp/log/pow).
# Newton solve with the structured factors: ~1450 flops; factorisation: ~2700 flops; per
# accepted step the Nordsieck map, correction update, error test and weights: ~1100 flops.
F_RHS = 1292
F_SETUP, F_NEWTON, F_STEP = 2700, 1450, 1100
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ord_bdf_population_10k_traffic.txt")
VALU_FILE = os.path.join(ROOT, "profiles", "cell_kernels_valu_counters.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_population(ncell: int):
    pop = cep.workload.ord_population(20251204, 50, 4, 50)
    # host-side scheduling: cells grouped by BCL so each wavefront holds equal-length beats
    order = np.argsort(pop["bcl"], kind="stable")
    if ncell < len(order):
        order = order[np.linspace(0, len(order) - 1, ncell).astype(int)]
    return {k: v[order] for k, v in pop.items()}


def make_batch(pop, prec, nbeats_total, nwin=None):
    n = len(pop["bcl"])
    b = cep.CellBatch(cep.MODEL_ORD_CIPA, n, prec)
    b.set_cells(celltype=pop["celltype"], conc=pop["conc"], herg=pop["herg"], gscale=pop["gscale"],
                bcl=pop["bcl"])
    b.set_run_config(dt=0.1, rtol=1e-6, atol=1e-8, maxord=5, nbeats_total=nbeats_total,
                     nwin=nbeats_total if nwin is None else nwin)
    b.init_consts()
    return b


def parse_group_cells(spec: str) -> dict:
    """'2000:8,1000:16' -> {2000.0: 8, 1000.0: 16} (cells per 16-slot wavefront by BCL)."""
    out = {}
    for part in filter(None, spec.split(",")):
        k, v = part.split(":")
        out[float(k)] = int(v)
    return out


def make_slot_batch(pop, prec, nbeats_total, per_group: dict):
    """The population in the BDF kernel's wavefront groups laid out by BCL class
    (workload.group_slots): padding slots parked.  Returns the batch and each cell's slot."""
    slots = cep.workload.group_slots(pop["bcl"], per_group)
    real = slots >= 0
    src = np.where(real, slots, slots[real][0])
    b = make_batch({k: v[src] for k, v in pop.items()}, prec, nbeats_total)
    if not real.all():
        b.park_cells(np.nonzero(~real)[0])
    pos = np.empty(len(pop["bcl"]), dtype=np.int64)
    pos[slots[real]] = np.nonzero(real)[0]
    return b, pos


def flops_of(stats: np.ndarray) -> float:
    """Counted flops of the integration work recorded in the per-cell statistics
    (steps, rhs evals incl. Jacobian columns, jacobians, setups, Newton iterations, ...,
    output points: one rhs each)."""
    nst, nfe, nje, nsetups, nni = (stats[:, k].astype(float) for k in range(5))
    nout = stats[:, 7].astype(float)
    return float(np.sum((nfe + nout) * F_RHS + nsetups * F_SETUP + nni * F_NEWTON + nst * F_STEP))


def read_profile(path):
    """The `key = value` lines of a committed rocprofv3 summary ('#' lines are commentary)."""
    kv = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            kv[k.strip()] = v.split("#")[0].strip()
    return kv


def read_traffic(path=TRAFFIC_FILE, keys=("hbm_bytes_per_beat_launch",)):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration): the sum of the named entries, or None."""
    try:
        kv = read_profile(path)
        return float(sum(float(kv[k]) for k in keys))
    except (OSError, KeyError, ValueError):
        return None


def read_valu(prefix: str, path=VALU_FILE):
    """VALU-busy, active-lane, issued FP32/FP64 and transcendental figures of one cell kernel from
    the committed rocprofv3 counter summary (profile-derived, not measured by this run), or None."""
    try:
        kv = read_profile(path)
        out = {k: (float(kv[f"{prefix}_{k}"]) if f"{prefix}_{k}" in kv else None)
               for k in ("valu_busy_pct", "valu_utilization_pct", "fp32_issued_tflops",
                         "fp64_issued_tflops", "trans_frac")}
        out["kernel"] = kv[f"{prefix}_kernel"]
        out["source"] = "rocprofv3 SQ counters, profiles/" + os.path.basename(path)
        return out
    except (OSError, KeyError, ValueError):
        return None


def host_info():
    """CPU model and thread counts of the host the CPU legs run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    return {"cpu_model": model, "nproc": nproc, "affinity": avail}


# Host threads of the CPU legs: one GPU's share of the box's host cores is 16, whatever nproc
# shows (the box is shared by GPUs).
CPU_THREADS = 16


def cpu_threads():
    return max(1, min(CPU_THREADS, host_info()["affinity"]))


def cpu_baseline(pop, states, seconds: float, beat_index: int):
    """The FP64 CPU oracle (test infrastructure) on a bounded sample of the same population:
    cells drawn at random (so the BCL mix is the population's), each started from the state the
    GPU batch held when the timed beats began (beat `beat_index`) and run for one beat, one cell
    stream per host thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = cpu_threads()
    n = len(pop["bcl"])
    rng = np.random.default_rng(7)

    def run(idx):
        cfgs = (O.OrdCellCfg * len(idx))()
        for k, i in enumerate(idx):
            cfgs[k] = O.make_ord_cfg(celltype=pop["celltype"][i], conc=pop["conc"][i], herg=pop["herg"][i],
                                     gscale=pop["gscale"][i], bcl=pop["bcl"][i], nbeats=1, nwin=1)
        res = (O.BmResult * len(idx))()
        # the oracle driver takes per-cell start states through its states_out rows when
        # started from given states: run each cell on its own thread slot
        S = np.ascontiguousarray(states[idx])
        wall = O.lib().oracle_ord_run_many_from(cfgs, len(idx), threads, O.ptr(S), res)
        return wall

    probe = rng.choice(n, size=4 * threads, replace=False)
    w0 = run(probe)
    rate = len(probe) / max(w0, 1e-3)
    ncell = int(min(n, max(4 * threads, rate * seconds)))
    idx = rng.choice(n, size=ncell, replace=False)
    wall = run(idx)
    info = host_info()
    return {"value": ncell / wall, "unit": "cell-beats/s", "cores": threads, "kind": "port",
            "sample": f"{ncell} random cells of the same population x 1 beat (beat {beat_index + 1}, "
                      f"from the states the GPU batch timed from), FP64 BDF oracle rtol 1e-6, "
                      f"{threads} threads, {wall:.1f} s",
            "cpu_model": info["cpu_model"], "nproc": info["nproc"]}


def stratified_cells(pop, per_bcl: int):
    """per_bcl cells of each BCL class, evenly spaced within the class."""
    sel = []
    for b in np.unique(pop["bcl"]):
        idx = np.nonzero(pop["bcl"] == b)[0]
        sel += list(idx[np.linspace(0, len(idx) - 1, min(per_bcl, len(idx))).astype(int)])
    return np.array(sel, dtype=int)


def oracle_runs(sub, nbeats: int, rtol: float, atol: float, threads: int, nwin=None):
    """The FP64 oracle on each cell of `sub`: biomarker rows, final states and last-beat Vm
    (biomarker window: the last `nwin` beats, all of them by default)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from concurrent.futures import ThreadPoolExecutor
    n = len(sub["bcl"])

    def one(k):
        cfg = O.make_ord_cfg(celltype=sub["celltype"][k], conc=sub["conc"][k], herg=sub["herg"][k],
                             gscale=sub["gscale"][k], bcl=sub["bcl"][k], nbeats=nbeats,
                             nwin=nbeats if nwin is None else nwin,
                             rtol=rtol, atol=atol)
        rc, S, res, vm, _ = O.ord_run(cfg, trace=True)
        return rc, S, [res[f] for f in cep.BM_FIELDS], vm

    with ThreadPoolExecutor(threads) as ex:      # ctypes releases the GIL inside the oracle
        r = list(ex.map(one, range(n)))
    L = max(len(x[3]) for x in r)
    vm = np.full((n, L), np.nan)
    for k, x in enumerate(r):
        vm[k, :len(x[3])] = x[3]
    return {"rc": np.array([x[0] for x in r]), "S": np.array([x[1] for x in r]),
            "bm": np.array([x[2] for x in r], dtype=float), "vm": vm}


def compare_to_oracle(bm, vm, o6, o7):
    """Errors of one GPU solver against the FP64 oracle at the same tolerance (o6), with the
    oracle's own change when its tolerance is tightened ten times (o7) as the resolution of each
    figure.  A cell 'repolarises' when both the solver and the oracle find an APD90 in the
    selected beat; a figure is 'resolved' when the oracle changes it by less than the bar
    (1 ms, 1 %) between the two tolerances."""
    bi = {k: i for i, k in enumerate(cep.BM_FIELDS)}
    a, q, p = bi["apd90"], bi["qnet"], bi["steepest_pace"]
    B, B7 = o6["bm"], o7["bm"]
    rep = (bm[:, a] > 0) & (B[:, a] > 0)
    ea = np.abs(bm[:, a] - B[:, a])
    ea7 = np.abs(B7[:, a] - B[:, a])
    den = np.maximum(np.abs(B[:, q]), 1e-300)
    eq = np.abs(bm[:, q] - B[:, q]) / den
    eq7 = np.abs(B7[:, q] - B[:, q]) / den
    qres = rep & (eq7 <= 0.01)
    ares = rep & (ea7 <= 1.0)
    vml = np.nanmax(np.abs(vm - o6["vm"]), axis=1)
    over = [{"cell": int(k), "qnet_oracle": float(B[k, q]), "rel_err": float(eq[k]),
             "oracle_self_rel": float(eq7[k]), "apd90_oracle": float(B[k, a])}
            for k in np.nonzero(rep & (eq > 0.01))[0]]
    fields = {f: float(np.max(np.abs(bm[rep, bi[f]] - B[rep, bi[f]]))) if rep.any() else None
              for f in cep.BM_FIELDS if f != "steepest_pace"}
    # Against the tighter oracle, a stand-in for the exact solution, beside the error of the
    # reference solver itself at the same tolerance (the oracle at rtol 1e-6 against 1e-7), on the
    # cells whose last beat the oracle resolves (its two tolerances end within 1 mV of each other)
    # and that repolarise in all three runs.  A solver is as accurate as the reference solver on a
    # cell when its error exceeds the reference solver's by no more than the bar (1 ms, 1 %).
    S = (np.nanmax(np.abs(o7["vm"] - o6["vm"]), axis=1) <= 1.0) & rep & (B7[:, a] > 0)
    ea_s, ea_r = np.abs(bm[:, a] - B7[:, a]), np.abs(B[:, a] - B7[:, a])
    q7 = np.maximum(np.abs(B7[:, q]), 1e-300)
    eq_s, eq_r = np.abs(bm[:, q] - B7[:, q]), np.abs(B[:, q] - B7[:, q])
    vm_s, vm_r = np.nanmax(np.abs(vm - o7["vm"]), axis=1), np.nanmax(np.abs(o6["vm"] - o7["vm"]), axis=1)
    beyond = S & ((ea_s > ea_r + 1.0) | (eq_s > eq_r + 0.01 * q7))
    tight = None
    if S.any():
        tight = {
            "cells": int(S.sum()),
            "apd90_max_abs_err_ms": float(ea_s[S].max()), "apd90_max_abs_err_ms_reference_solver": float(ea_r[S].max()),
            "qnet_median_rel_err": float(np.median(eq_s[S] / q7[S])),
            "qnet_median_rel_err_reference_solver": float(np.median(eq_r[S] / q7[S])),
            "qnet_max_rel_err": float((eq_s[S] / q7[S]).max()),
            "qnet_max_rel_err_reference_solver": float((eq_r[S] / q7[S]).max()),
            "vm_linf_mv_last_beat_median": float(np.median(vm_s[S])),
            "vm_linf_mv_last_beat_median_reference_solver": float(np.median(vm_r[S])),
            "cells_beyond_reference_error_plus_bar": [
                {"cell": int(k), "qnet": float(B7[k, q]), "qnet_abs_err": float(eq_s[k]),
                 "qnet_abs_err_reference_solver": float(eq_r[k]), "apd90_err_ms": float(ea_s[k]),
                 "apd90_err_ms_reference_solver": float(ea_r[k])} for k in np.nonzero(beyond)[0]],
        }
    return {
        "cells_repolarising": int(rep.sum()),
        "apd90_max_abs_err_ms": float(ea[rep].max()) if rep.any() else None,
        "apd90_frac_within_1ms": float(np.mean(ea[rep] <= 1.0)) if rep.any() else None,
        "apd90_max_abs_err_ms_resolved": float(ea[ares].max()) if ares.any() else None,
        "qnet_max_rel_err": float(eq[rep].max()) if rep.any() else None,
        "qnet_frac_within_1pct": float(np.mean(eq[rep] <= 0.01)) if rep.any() else None,
        "qnet_max_rel_err_resolved": float(eq[qres].max()) if qres.any() else None,
        "qnet_cells_resolved": int(qres.sum()),
        "qnet_cells_over_1pct": over,
        "steepest_pace_agree_frac": float(np.mean(bm[:, p] == B[:, p])),
        "vm_linf_mv_last_beat": {"median": float(np.median(vml)), "p90": float(np.percentile(vml, 90)),
                                 "max_repolarising": float(vml[rep].max()) if rep.any() else None,
                                 "max": float(vml.max())},
        "max_abs_err_by_field_repolarising": fields,
        "vs_oracle_rtol1e-7_beside_reference_solver": tight,
    }


def accuracy_check(pop, bm_timed, per_bcl: int, nbeats: int, window=None, dump=None):
    """The timed mixed solver and the plain per-job FP64 GPU solver against the FP64 CPU oracle
    on BCL-stratified cells after the beats the bench ran.  `window`: the biomarkers of the
    last `window` beats (the selected beat among them), all beats by default; 1 compares the
    last beat itself, with no selection between nearly equal beats.  `dump`: an .npz file for
    the per-cell figures."""
    sel = stratified_cells(pop, per_bcl)
    sub = {k: v[sel] for k, v in pop.items()}
    threads = cpu_threads()
    from concurrent.futures import ThreadPoolExecutor
    ex = ThreadPoolExecutor(1)
    # the oracle runs on the host while the GPU runs the two solvers
    fut = ex.submit(lambda: (oracle_runs(sub, nbeats, 1e-6, 1e-8, threads, window),
                             oracle_runs(sub, nbeats, 1e-7, 1e-9, threads, window)))
    out = {}
    for prec, tag in ((cep.PREC_MIXED, "mixed"), (cep.PREC_FP64, "fp64")):
        log(f"bench: {tag} GPU job of {len(sel)} check cells, {nbeats} beats")
        b = make_batch(sub, prec, 200) if window is None else make_batch(sub, prec, nbeats, window)
        b.enable_vm_trace(True)
        t0 = time.perf_counter()
        for done in range(0, nbeats, 25):         # launches of 25 beats (bitwise as one launch)
            b.run_beats(min(25, nbeats - done))
            if nbeats > 25:
                log(f"bench: {tag} check job: {min(done + 25, nbeats)} of {nbeats} beats")
        t1 = time.perf_counter()
        out[tag] = {"bm": b.biomarker_array(), "vm": b.vm_trace(), "rate": len(sel) * nbeats / (t1 - t0),
                    "failed": int(np.sum(b.status() != 0))}
        b.close()
    o6, o7 = fut.result()
    ex.shutdown()
    if dump:
        np.savez(dump, cells=sel, bcl=sub["bcl"], fields=np.array(cep.BM_FIELDS), bm_mixed=out["mixed"]["bm"],
                 bm_fp64=out["fp64"]["bm"], vm_mixed=out["mixed"]["vm"], vm_fp64=out["fp64"]["vm"],
                 bm_o6=o6["bm"], bm_o7=o7["bm"], vm_o6=o6["vm"], vm_o7=o7["vm"], rc_o6=o6["rc"], rc_o7=o7["rc"])
    mx = compare_to_oracle(out["mixed"]["bm"], out["mixed"]["vm"], o6, o7)
    f64 = compare_to_oracle(out["fp64"]["bm"], out["fp64"]["vm"], o6, o7)
    # the oracle against itself at the tighter tolerance: the resolution of the reference figures
    selfc = compare_to_oracle(o7["bm"], o7["vm"], o6, o7)
    # the plain mixed job against the timed batch, when both ran the same beats
    same = bool(np.array_equal(out["mixed"]["bm"], bm_timed[sel])) if bm_timed is not None else None
    gate = {
        "apd90_le_1ms_every_repolarising_cell": bool(mx["apd90_max_abs_err_ms"] is not None
                                                     and mx["apd90_max_abs_err_ms"] <= 1.0),
        "qnet_le_1pct_every_resolved_cell": bool(mx["qnet_max_rel_err_resolved"] is not None
                                                 and mx["qnet_max_rel_err_resolved"] <= 0.01),
    }
    gate["pass"] = all(gate.values())
    return {
        "oracle": "FP64 CPU oracle (CVODE-style BDF restated), rtol 1e-6/atol 1e-8 as the GPU; "
                  "rtol 1e-7/atol 1e-9 for the oracle's own resolution",
        "cells": int(len(sel)), "cells_per_bcl": per_bcl, "beats": nbeats,
        "window_beats": nbeats if window is None else window,
        "mixed_job_equals_timed_batch": same,
        "mixed_vs_oracle": mx,
        "fp64_gpu_vs_oracle": f64,
        "oracle_rtol1e-7_vs_rtol1e-6": {k: selfc[k] for k in ("apd90_max_abs_err_ms", "qnet_max_rel_err",
                                                               "steepest_pace_agree_frac", "vm_linf_mv_last_beat")},
        "gate": gate,
        "fp64_plain_job_cell_beats_per_s": out["fp64"]["rate"],
        "oracle_failed_cells": int(np.sum(o6["rc"] != 0)),
    }


def serving_latency(njobs: int, rate: float, beats):
    """Serving session: Poisson arrivals of screening jobs (1-400 cells, ORd-CiPA or TT06,
    BCL 500/1000/2000 ms, `beats` beats each) into the cell-level continuous-batching loop."""
    jobs = serving.synthetic_jobs(20251206, njobs, beats=beats)
    srv = serving.Server(mem_budget=32 << 30)
    return _serve(srv, jobs, njobs, rate, beats)


def _serve(srv, jobs, njobs: int, rate: float, beats):
    fin, wall = serving.run_poisson(srv, jobs, rate, seed=1, timeout=900)
    launches = {f"model{k[0]}_bcl{int(k[2])}": p.launches for k, p in srv.pools.items()}
    pool_cells = {f"model{k[0]}_bcl{int(k[2])}": p.capacity for k, p in srv.pools.items()}
    srv.close()
    p = serving.latency_percentiles(fin)
    cb = sum(j.ncells * j.nbeats for j in fin)
    bad = int(sum(int(np.sum(j.status != 0)) for j in fin))
    return {"p50": p["p50"], "p99": p["p99"], "jobs": len(fin), "rate_jobs_per_s": rate,
            "cell_beats_per_s": cb / wall, "failed_cells": bad, "wall_s": wall,
            "launches_per_pool": launches, "cells_per_pool": pool_cells,
            "workload": f"{njobs} jobs, Poisson {rate}/s, cells log-uniform 1-400, {beats[0]}-{beats[1]} beats, "
                        "ORd-CiPA or TT06 (p=0.5), BCL 500/1000/2000 ms per job, mixed precision, "
                        "one resident batch per model and BCL shared at the cell level"}


# Tissue (monodomain) workload: 512 x 512 Q1 elements, 0.1 mm, one seeded drug at 2 x Cmax.
TISSUE_NX, TISSUE_H, TISSUE_DT = 512, 0.1, 0.01
# algorithmic HBM bytes per node and CG iteration (FP64): the SpMV with the direction update
# fused, 9 coefficients + z + previous p read, new p + q written = 104; the vector update, x r
# read+write, p q dinv read, z written = 64
B_SPMV, B_UPDATE = 104, 64
PEAK_HBM = 8.0e12
TISSUE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "tissue_monodomain_513x513_traffic.txt")


def tissue_setup(nx: int, precision: int):
    """The 2-D tissue of the monodomain workload: fibre angle linear 0 -> 60 deg across x,
    sigma_l / sigma_t = 0.12 / 0.03 S/m as D = sigma / (chi Cm) with chi = 1400 /cm and
    Cm = 1 uF/cm^2 (0.0857 / 0.0214 mm^2/ms), ORd-CiPA endo cells with one seeded drug at
    2 x Cmax, 5 x 5 paced corner nodes."""
    rng = np.random.default_rng(20251207)
    herg, cmax = cep.workload.draw_drug(rng)
    tis = cep.Tissue(nx, nx, TISSUE_H, TISSUE_H, precision)
    tis.set_cells(celltype=0.0, conc=2.0 * cmax, herg=herg)
    dl, dtr = 0.12 / 1.4, 0.03 / 1.4
    th = np.deg2rad(60.0) * np.tile((np.arange(nx) + 0.5) / nx, nx)
    c, s = np.cos(th), np.sin(th)
    sxx, sxy, syy = dl * c * c + dtr * s * s, (dl - dtr) * c * s, dl * s * s + dtr * c * c
    tis.set_conductivity(sxx, sxy, syy)
    tis.set_stimulus([tis.node(i, j) for i in range(5) for j in range(5)], amp_scale=2.0, bcl=1000.0,
                     start=0.0, duration=1.0)
    return tis, (sxx, sxy, syy), (herg, 2.0 * cmax)


def tissue_cpu_baseline(tis, tensors, drug, nx: int):
    """The FP64 oracle on the host, one thread: one diffusion solve of the current state with
    the oracle's Jacobi-PCG on the assembled CSR system, and the oracle's right-hand side on a
    sample of the nodes' states (Strang: two reaction half-steps per step).  The reaction's
    state updates are not timed, so this slightly favours the host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    import tissue_oracle as T
    t0 = time.perf_counter()
    Acsr, Bcsr = T.assemble_csr(nx, nx, TISSUE_H, TISSUE_H, *tensors, TISSUE_DT)
    v = tis.vm()
    rp, col, val = Bcsr
    b = np.add.reduceat(val * v[col], rp[:-1])
    t1 = time.perf_counter()
    _, it = T.cg_csr(Acsr, b, 1e-5)
    t_cg = time.perf_counter() - t1
    st = tis.states()
    nn = st.shape[0]
    ns = min(nn, 20000)
    sel = np.linspace(0, nn - 1, ns).astype(int)
    Cc, _ = O.ord_init(0.0, drug[1], drug[0])
    Crow = np.ascontiguousarray(np.broadcast_to(Cc, (ns, len(Cc))))
    S = np.ascontiguousarray(st[sel])
    R = np.zeros_like(S)
    t_rhs = O.lib().oracle_ord_rates_many(ns, O.ptr(Crow), O.ptr(S), O.ptr(R), tis.time) / ns
    per_step = t_cg + 2.0 * nn * t_rhs
    return {"value": nn / per_step, "unit": "node-steps/s", "cores": 1, "kind": "port",
            "sample": f"one oracle Jacobi-PCG solve of the {nn}-node system ({it} iterations, "
                      f"{t_cg:.2f} s) + oracle RHS on {ns} node states x 2 half-steps, scaled to "
                      f"{nn} nodes; assembly {t1 - t0:.1f} s not counted"}


# TT06 Rush-Larsen: counted by oracle/opcount.cpp, flops per step = RHS 460 + gate/state update
# 74 (63 exp, 4 log and 27 pow per step are not counted)
F_TT_STEP = 460 + 74


def tt06_main(args):
    """ten Tusscher 2006 population, one lane per cell, fixed-dt Rush-Larsen."""
    prec = cep.PREC_MIXED if args.precision == "mixed" else cep.PREC_FP64
    K = args.steps if args.steps is not None else 3
    W = args.warmup if args.warmup is not None else 1
    n = args.tt_cells
    dt = 0.02
    pop = cep.workload.tt_population(20251205, n)
    herg = np.concatenate([pop["ic50"], pop["hill"]], axis=1)

    def batch(sel, pr):
        b = cep.CellBatch(cep.MODEL_TT06, len(sel), pr)
        b.set_cells(celltype=pop["celltype"][sel], conc=pop["conc"][sel], herg=herg[sel],
                    gscale=pop["gscale"][sel], bcl=pop["bcl"][sel])
        b.set_run_config(dt=dt, nbeats_total=500, nwin=500)
        b.init_consts()
        return b

    allc = np.arange(n)
    b = batch(allc, prec)
    log(f"bench: TT06 {n} cells, warmup {W} beat(s)")
    if W:
        b.run_beats(W)
    b.kernel_time()
    t0 = time.perf_counter()
    for _ in range(K):
        b.run_beats(1)
    t1 = time.perf_counter()
    nl, kavg_ms, _ = b.kernel_time()
    bm = b.biomarker_array()
    bad = int(np.sum(b.status() != 0))
    b.close()
    elapsed = t1 - t0
    value = n * K / elapsed
    steps_per_beat = int(round(1000.0 / dt))
    flops_launch = n * steps_per_beat * F_TT_STEP
    achieved = flops_launch / (kavg_ms * 1e-3) if kavg_ms > 0 else 0.0
    peak = PEAK_FP32 if prec == cep.PREC_MIXED else PEAK_FP64
    chk = {}
    if args.check_cells > 0 and prec == cep.PREC_MIXED:
        sel = np.linspace(0, n - 1, args.check_cells).astype(int)
        bi = {k: i for i, k in enumerate(cep.BM_FIELDS)}
        vm_linf = None
        if args.check_beats > 0:
            # a run of its own: the sample for check_beats beats in both precisions (the window
            # covers every beat, so both select their steepest beat among all of them)
            ends = {}
            for pr in (cep.PREC_MIXED, cep.PREC_FP64):
                bb = batch(sel, pr)
                for done in range(0, args.check_beats, 25):
                    bb.run_beats(min(25, args.check_beats - done))
                    log(f"bench: TT06 check, {'mixed' if pr == cep.PREC_MIXED else 'fp64'}: "
                        f"{min(done + 25, args.check_beats)} of {args.check_beats} beats")
                ends[pr] = (bb.biomarker_array(), bb.states()[:, 0].copy(), bb.status().copy())
                bb.close()
            (bmc, vmc, stc), (bm64, vm64, st64) = ends[cep.PREC_MIXED], ends[cep.PREC_FP64]
            vm_linf = float(np.abs(vmc - vm64).max())
            bad_check = int(np.sum(stc != 0) + np.sum(st64 != 0))
        else:
            b64 = batch(sel, cep.PREC_FP64)
            b64.run_beats(W + K)
            bm64 = b64.biomarker_array()
            b64.close()
            bmc = bm[sel]
            bad_check = None
        ea = np.abs(bmc[:, bi["apd90"]] - bm64[:, bi["apd90"]])
        chk = {"apd90_max_abs_err_ms": float(ea.max()), "apd90_median_abs_err_ms": float(np.median(ea)),
               "apd90_frac_within_1ms": float(np.mean(ea <= 1.0)),
               "cad90_max_abs_err_ms": float(np.abs(bmc[:, bi["cad90"]] - bm64[:, bi["cad90"]]).max()),
               "cells_checked": int(len(sel)), "beats": int(args.check_beats or (W + K))}
        if vm_linf is not None:
            chk.update({"vm_linf_mv_at_end": vm_linf, "failed_cells": bad_check})
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        threads = cpu_threads()
        ncpu = max(threads, int(args.cpu_seconds * threads / 0.12))
        sel = np.linspace(0, n - 1, ncpu).astype(int)
        cfgs = (O.TtCellCfg * ncpu)()
        for k, i in enumerate(sel):
            cfgs[k] = O.make_tt_cfg(celltype=pop["celltype"][i], conc=pop["conc"][i], ic50=pop["ic50"][i],
                                    hill=pop["hill"][i], gscale=pop["gscale"][i], bcl=pop["bcl"][i], nbeats=1,
                                    nwin=1, dt=dt)
        res = (O.BmResult * ncpu)()
        log(f"bench: CPU baseline, {ncpu} cells x 1 beat")
        wall = O.lib().oracle_tt_run_many(cfgs, ncpu, threads, res, None)
        cpu = {"value": ncpu / wall, "unit": "cell-beats/s", "cores": threads, "kind": "port",
               "sample": f"{ncpu} cells x 1 beat of the same population, FP64 Rush-Larsen oracle, "
                         f"{threads} threads, {wall:.1f} s",
               "cpu_model": host_info()["cpu_model"], "nproc": host_info()["nproc"]}
    print(json.dumps({
        "metric": "TT06 cell-beats/s on 1 GPU (Rush-Larsen dt 0.02 ms)",
        "value": value, "unit": "cell-beats/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * elapsed / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-rates/f64-state" if prec == cep.PREC_MIXED else "f64",
        "data": "synthetic (seeded drug/IC50/conductance population)",
        "config": {"workload": f"ten Tusscher 2006 population, {n} cells endo/epi/M, Hill block of "
                               "IKr/ICaL/INa at 1-4 x Cmax, 10 lognormal(0, 0.25) conductance scales, BCL "
                               "1000 ms, Rush-Larsen dt 0.02 ms, step = 1 beat", "cells": n},
        "failed_cells": bad,
        "accuracy_vs_fp64": chk,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel": f"tt_rl_kernel<{'float' if prec == cep.PREC_MIXED else 'double'}>",
                     "kernel_avg_ms": kavg_ms, "flops_per_launch": flops_launch,
                     "flops_note": "534 counted flops per cell-step; 63 exp, 4 log, 27 pow per step not counted",
                     "valu_counters": read_valu("tt") if prec == cep.PREC_MIXED else None},
        "cpu_baseline": cpu,
    }), flush=True)


def tune_main(args):
    """The launch-shape autotuner on each model at a small and a production cell count: every
    candidate's time for one beat of the batch's own cells, and the choice.  A tuning cache
    (--tune-cache) keeps the choices for later runs."""
    G = os.path.join(ROOT, "tests", "golden")
    cases = []

    def ord_cipa(n):
        pop = cep.workload.ord_population(20251204, 50, 4, max(1, n // 200))
        order = np.argsort(pop["bcl"], kind="stable")
        pop = {k: v[order][:n] for k, v in pop.items()}
        b = make_batch(pop, cep.PREC_MIXED, 200)
        return b, f"ORd-CiPA BDF, {n} cells, BCL 500/1000/2000 stored shortest first, mixed"

    def tt06(n):
        pop = cep.workload.tt_population(20251205, n)
        b = cep.CellBatch(cep.MODEL_TT06, n, cep.PREC_MIXED)
        b.set_cells(celltype=pop["celltype"], conc=pop["conc"],
                    herg=np.concatenate([pop["ic50"], pop["hill"]], axis=1), gscale=pop["gscale"], bcl=pop["bcl"])
        b.set_run_config(dt=0.02, nbeats_total=500, nwin=500)
        b.init_consts()
        return b, f"TT06 Rush-Larsen, {n} cells, mixed"

    def ord11(n):
        q = np.loadtxt(os.path.join(G, "quinidine_IC50_samples10.csv"), delimiter=",", skiprows=1)
        s0 = np.loadtxt(os.path.join(G, "ord2011_output_orudy.dat"))
        conc = np.array([(3237.0, 6474.0, 9711.0)[i % 3] for i in range(n)])
        b = cep.CellBatch(cep.MODEL_ORD2011, n, cep.PREC_FP64)
        b.set_cells(celltype=np.full(n, 2.0), conc=conc, herg=np.vstack([q[i % len(q)] for i in range(n)]),
                    bcl=np.full(n, 2000.0))
        b.set_run_config(dt=1.0, nbeats_total=1000, nwin=250)
        b.init_consts()
        b.set_states(np.repeat(s0[None, :], n, axis=0))
        return b, f"ORd 2011 Rush-Larsen, {n} quinidine cells, FP64"

    plan = [(ord_cipa, 10000), (ord_cipa, 40000), (tt06, 4000), (tt06, 200000), (ord11, 30), (ord11, 100000)]
    for make, n in plan:
        b, name = make(n)
        default = b.launch_shape()
        log(f"bench: tuning {name}")
        rows, best = b.autotune(1, args.tune_cache)
        chosen = b.launch_shape()
        b.close()
        dms = [r["ms"] for r in rows if all(r[k] == default[k] for k in cep.SHAPE_FIELDS)]
        cases.append({"workload": name, "cells": n, "candidates": rows, "chosen": chosen, "default": default,
                      "default_ms": dms[0] if dms else None, "chosen_ms": rows[best]["ms"],
                      "speedup_vs_default": (dms[0] / rows[best]["ms"]) if dms else None})
    # the population's wavefront layout by BCL class (host side: cells per 16-slot group of the
    # BDF kernel, padding parked): one warm beat, then one timed beat per candidate; results
    # are bitwise those of the 16-cell layout
    pop = make_population(10000)
    lay = []
    ref = None
    for spec in ("", "2000:12", "2000:8", "2000:7", "2000:6", "2000:8,1000:12"):
        pg = parse_group_cells(spec)
        b, pos = make_slot_batch(pop, cep.PREC_MIXED, 200, pg)
        b.run_beats(1)
        t0 = time.perf_counter()
        b.run_beats(1)
        ms = 1e3 * (time.perf_counter() - t0)
        bm = b.biomarker_array()[pos]
        same = True if ref is None else bool(np.array_equal(bm, ref))
        ref = bm if ref is None else ref
        lay.append({"cells_per_wavefront_by_bcl": spec or "16 everywhere", "slots": b.n, "ms": ms,
                    "bitwise_equal_to_16": same})
        log(f"bench: layout '{spec or '16'}': {b.n} slots, {ms:.1f} ms per beat")
        b.close()
    best = min(lay, key=lambda r: r["ms"])
    cases.append({"workload": "ORd-CiPA BDF, 10,000-cell population, wavefront layout by BCL class",
                  "cells": 10000, "candidates": lay, "chosen": best["cells_per_wavefront_by_bcl"],
                  "default_ms": lay[0]["ms"], "chosen_ms": best["ms"],
                  "speedup_vs_default": lay[0]["ms"] / best["ms"]})
    print(json.dumps({"metric": "launch-shape autotuner: one-beat trial per candidate", "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "cases": cases}), flush=True)


def serving_main(args):
    """Arrival-rate sweep of the serving loop: for each rate a fresh server and the same
    seeded jobs (ORd-CiPA or TT06, 1-400 cells, 2-4 beats), submitted with Poisson gaps."""
    rates = [float(r) for r in args.serve_rates.split(",")]
    beats = tuple(int(x) for x in args.serve_sweep_beats.split(","))
    sweep = []
    for rate in rates:
        log(f"bench: serving at {rate} jobs/s, {args.serve_jobs} jobs")
        jobs = serving.synthetic_jobs(20251206, args.serve_jobs, beats=beats)
        srv = serving.Server(mem_budget=32 << 30)
        r = _serve(srv, jobs, args.serve_jobs, rate, beats)
        log(f"bench: rate {rate}: p50 {r['p50']:.0f} ms, p99 {r['p99']:.0f} ms, "
            f"{r['cell_beats_per_s']:.0f} cell-beats/s")
        sweep.append({k: r[k] for k in ("rate_jobs_per_s", "p50", "p99", "cell_beats_per_s", "jobs",
                                        "failed_cells", "wall_s", "launches_per_pool")})
    best = max(sweep, key=lambda e: e["cell_beats_per_s"])
    print(json.dumps({
        "metric": "ORd cell-beats/s on 1 GPU at APD90 err <=1 ms; p50/p99 job latency",
        "value": best["cell_beats_per_s"], "unit": "cell-beats/s", "n_gpus": 1,
        "steps": len(rates), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-rates/f64-state",
        "data": "synthetic (seeded screening jobs)",
        "config": {"workload": f"serving sweep: {args.serve_jobs} jobs per rate, Poisson arrivals, cells "
                               f"log-uniform 1-400, {beats[0]}-{beats[1]} beats, ORd-CiPA or TT06 (p=0.5), "
                               "BCL 500/1000/2000 ms per job, one resident batch per model and BCL "
                               "shared at the cell level"},
        "latency_ms": {"p50": best["p50"], "p99": best["p99"], "rate_jobs_per_s": best["rate_jobs_per_s"]},
        "sweep": sweep,
    }), flush=True)


def tissue_main(args):
    K = args.steps if args.steps is not None else 2000
    W = args.warmup if args.warmup is not None else 200
    nx = args.tissue_nx
    prec = cep.PREC_MIXED if args.precision == "mixed" else cep.PREC_FP64
    tis, tensors, drug = tissue_setup(nx, prec)
    nn = tis.nn
    log(f"bench: tissue {nx}x{nx} elements ({nn} nodes), warmup {W} steps")
    if W:
        tis.step(W, TISSUE_DT, cep.SPLIT_STRANG)
    tis.kernel_times()
    log(f"bench: timing {K} steps")
    t0 = time.perf_counter()
    iters = tis.step(K, TISSUE_DT, cep.SPLIT_STRANG)
    t1 = time.perf_counter()
    avg, n = tis.kernel_times()
    vm = tis.vm()
    elapsed = t1 - t0
    ipi = iters / K
    spmv_bytes = nn * B_SPMV
    ach = spmv_bytes / (avg[1] * 1e-3) if avg[1] > 0 else 0.0
    cpu = None if args.no_cpu_baseline else tissue_cpu_baseline(tis, tensors, drug, nx)
    tis.close()
    # FP32 rates: the same W + K steps in FP64 for the Vm difference they make
    acc = None
    if prec == cep.PREC_MIXED:
        log(f"bench: FP64 tissue for the same {W + K} steps")
        t64, _, _ = tissue_setup(nx, cep.PREC_FP64)
        t64.step(W + K, TISSUE_DT, cep.SPLIT_STRANG)
        d = np.abs(vm - t64.vm())
        t64.close()
        acc = {"vm_linf_mv_vs_fp64": float(d.max()), "vm_median_abs_mv_vs_fp64": float(np.median(d)),
               "steps": W + K, "simulated_ms": (W + K) * TISSUE_DT}
    out = {
        "metric": "tissue node-steps/s (2-D monodomain, ORd-CiPA nodes, Q1 FEM + Jacobi-PCG)",
        "value": nn * K / elapsed,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == cep.PREC_FP64 else "f32-rates/f64-state",
        "data": "synthetic (seeded drug, fibre-rotated conductivity)",
        "config": {"workload": f"{nx}x{nx} Q1 elements ({nn} nodes), dx 0.1 mm, sigma 0.12/0.03 S/m, "
                               "fibre angle 0->60 deg across x, 1 drug at 2xCmax, 5x5 corner stimulus, "
                               "dt 0.01 ms, Strang splitting, CG rtol 1e-5",
                   "nodes": nn},
        "cg_iters_per_step": ipi,
        "vm_range_mv": [float(np.min(vm)), float(np.max(vm))],
        "kernels_avg_ms": {"tissue_react": float(avg[0]), "cg_spmv": float(avg[1]), "cg_update": float(avg[2]),
                           "launches": [int(x) for x in n]},
        "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": ach / PEAK_HBM,
                     "traffic": read_traffic(TISSUE_TRAFFIC_FILE, ("cg_spmv.hbm_read_bytes", "cg_spmv.hbm_write_bytes"))
                     if nx == TISSUE_NX else None,
                     "kernel": "cg_spmv",
                     "bytes_per_launch": spmv_bytes,
                     "update_gbs": nn * B_UPDATE / (avg[2] * 1e-3) / 1e9 if avg[2] > 0 else None},
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": acc,
    }
    print(json.dumps(out), flush=True)


# ORd 2011 workload with a published reference time: the DrugSimulation quinidine run (10 IC50
# samples x 3 concentrations, 1000 paces at BCL 2000 ms from the shipped steady state, M cell,
# no Dutta) took 1048.7 s on 10 MPI ranks = 28.6 cell-beats/s.
ORD11_REF_CELL_BEATS_PER_S = 30 * 1000 / 1048.7
ORD11_EAD_LOGGED = {(8, 6474.0), (4, 9711.0), (9, 9711.0), (7, 9711.0), (0, 9711.0), (8, 9711.0)}


def ord11_inputs():
    """The reference's fixtures of its ORd 2011 run: quinidine IC50 samples and steady state."""
    G = os.path.join(ROOT, "tests", "golden")
    q = np.loadtxt(os.path.join(G, "quinidine_IC50_samples10.csv"), delimiter=",", skiprows=1)
    s0 = np.loadtxt(os.path.join(G, "ord2011_output_orudy.dat"))
    return q, s0


def ord11_batch(ncell, conc, ic50, prec, nbeats, s0):
    """ORd 2011 M cells at BCL 2000 ms from the shipped steady state, as the reference's run."""
    b = cep.CellBatch(cep.MODEL_ORD2011, ncell, prec)
    b.set_cells(celltype=np.full(ncell, 2.0), conc=conc, herg=ic50, bcl=np.full(ncell, 2000.0))
    b.set_run_config(dt=1.0, nbeats_total=nbeats, nwin=min(250, nbeats))
    b.init_consts()
    b.set_states(np.repeat(s0[None, :], ncell, axis=0))
    return b


def ord11_main(args):
    """The reference's own ORd 2011 run, on one GPU, plus a batched population of the same
    model for throughput at scale."""
    q, s0 = ord11_inputs()
    nbeats = args.steps if args.steps is not None else 1000
    keys = [(i, c) for c in (3237.0, 6474.0, 9711.0) for i in range(len(q))]
    n = len(keys)
    prec = cep.PREC_FP64 if (args.precision or "fp64") == "fp64" else cep.PREC_MIXED

    def batch(ncell, conc, ic50):
        return ord11_batch(ncell, conc, ic50, prec, nbeats, s0)

    log(f"bench: ORd 2011 quinidine run, {n} cells x {nbeats} beats")
    b = batch(n, np.array([c for _, c in keys]), np.vstack([q[i] for i, _ in keys]))
    b.kernel_time()
    t0 = time.perf_counter()
    b.run_beats(nbeats)
    t1 = time.perf_counter()
    nl, kavg, ktot = b.kernel_time()
    rec = b.biomarker_array()
    steps = b.stats()[:, 0].astype(float)
    bad = int(np.sum(b.status() != 0))
    b.close()
    ead = {k for k, r in zip(keys, rec) if r[cep.BM11_FIELDS.index("dvmdt_repol")] > 0}
    wall = t1 - t0
    value = n * nbeats / wall
    # batched population of the same model: the quinidine samples tiled over conc 0..4 x 9711 nM
    npop = args.ncell if args.ncell != 10000 else 100000
    popbeats = 2
    rng = np.random.default_rng(20251208)
    conc = rng.uniform(0.0, 4 * 9711.0, npop)
    ic50 = q[rng.integers(0, len(q), npop)]
    log(f"bench: ORd 2011 population, {npop} cells x {popbeats} beats")
    bp = batch(npop, conc, ic50)
    bp.kernel_time()
    t2 = time.perf_counter()
    bp.run_beats(popbeats)
    t3 = time.perf_counter()
    pl, pavg, _ = bp.kernel_time()
    psteps = float(bp.stats()[:, 0].sum())
    bp.close()
    # algorithmic work per step (oracle/opcount.cpp on the ORd 2011 restatement): RHS 1046
    # flops + Rush-Larsen/Euler update 146 flops (103 exp/log/pow apart)
    f_step = 1046 + 146
    ach = f_step * psteps / (pavg * pl * 1e-3) if pavg > 0 else 0.0
    peak = PEAK_FP64 if prec == cep.PREC_FP64 else PEAK_FP32
    out = {
        "metric": "ORd 2011 cell-beats/s (DrugSimulation quinidine run: 10 samples x 3 conc x 1000 beats)",
        "value": value,
        "unit": "cell-beats/s",
        "n_gpus": 1,
        "steps": nbeats,
        "warmup": 0,
        "ms_per_step": 1e3 * wall / nbeats,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / ORD11_REF_CELL_BEATS_PER_S if nbeats == 1000 else None,
        "dtype": "f64" if prec == cep.PREC_FP64 else "f32-rates/f64-state",
        "data": "reference fixtures (quinidine IC50 samples, steady state) + seeded population",
        "config": {"workload": "ORd 2011 M cell, quinidine IC50_samples10 x {3237, 6474, 9711} nM, BCL 2000, "
                               "adaptive Rush-Larsen 0.005-1 ms, features over the last 250 beats",
                   "cells": n, "beats": nbeats},
        "reference_wall_s": 1048.7,
        "wall_s": wall,
        "ead_cases": sorted([list(k) for k in ead]),
        "ead_matches_reference_log": ead == ORD11_EAD_LOGGED if nbeats == 1000 else None,
        "failed_cells": bad,
        "steps_per_cell_beat": float(steps.mean() / nbeats),
        "population": {"cells": npop, "beats": popbeats, "cell_beats_per_s": npop * popbeats / (t3 - t2),
                       "kernel_avg_ms": pavg},
        "roofline": {"bound": "valu", "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": ach / peak, "traffic": None, "kernel": "ord11_rl_kernel (population)",
                     "flops_per_step": f_step,
                     "valu_counters": read_valu("o11") if prec == cep.PREC_FP64 else None},
        "cpu_baseline": None if args.no_cpu_baseline else ord11_cpu_baseline(q, s0),
    }
    print(json.dumps(out), flush=True)


def ord11_cpu_baseline(q, s0):
    """The FP64 oracle on the host: the same quinidine cells, 20 beats each, 16 threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cfgs = [O.make_ord11_cfg(conc=c, ic50=q[i], bcl=2000.0, nbeats=20, nwin=20, celltype=2.0)
            for c in (3237.0, 6474.0, 9711.0) for i in range(len(q))]
    threads = cpu_threads()
    _, _, wall = O.ord11_run_many(cfgs, s0, nthreads=threads)
    return {"value": len(cfgs) * 20 / wall, "unit": "cell-beats/s", "cores": threads, "kind": "port",
            "sample": f"{len(cfgs)} cells x 20 beats of the same run, FP64 oracle, {threads} threads, {wall:.1f} s",
            "cpu_model": host_info()["cpu_model"], "nproc": host_info()["nproc"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("population", "tissue", "ord2011", "serving", "tt06", "tune"),
                    default="population")
    ap.add_argument("--tt-cells", type=int, default=200000)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--tissue-nx", type=int, default=TISSUE_NX)
    ap.add_argument("--ncell", type=int, default=10000)
    ap.add_argument("--group-cells", default="2000:7",
                    help="population: cells per 16-slot BDF wavefront by BCL ('bcl:cells,...'; '' = 16 everywhere)")
    ap.add_argument("--precision", choices=("mixed", "fp64"), default=None,
                    help="default mixed (FP32 rates; FP64 state, BDF/LU, CG)")
    ap.add_argument("--check-cells", type=int, default=64, help="tt06: FP64 check cells")
    ap.add_argument("--check-window", type=int, default=0,
                    help="population: biomarker window of the check jobs in beats (0: every beat run; "
                         "1: the last beat, no selection between beats)")
    ap.add_argument("--check-dump", default=None, help="population: .npz file for the check's per-cell figures")
    ap.add_argument("--check-beats", type=int, default=0,
                    help="population, tt06: run the check cells for this many beats "
                         "(0: the beats the timed batch ran)")
    ap.add_argument("--check-per-bcl", type=int, default=32,
                    help="population: oracle-checked cells per BCL class (0: no accuracy check)")
    ap.add_argument("--fp64-beats", type=int, default=2,
                    help="population: beats of the other precision on the full batch (0: skip)")
    ap.add_argument("--legs", type=int, default=1, help="population: compact TT06 and tissue legs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=120)
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serve-beats", default="10,40", help="population: beats per serving job (low,high)")
    ap.add_argument("--serve-rates", default="2,8,32", help="arrival rates of --workload serving")
    ap.add_argument("--serve-sweep-beats", default="10,100", help="--workload serving: beats per job (low,high)")
    ap.add_argument("--tune-cache", default=None, help="--workload tune: tuning cache file to write")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: run with --gpus 1 (cells shard trivially "
                          "across GPUs; see the design notes)"}))
        sys.exit(2)
    if args.workload == "ord2011":
        ord11_main(args)
        return
    if args.workload == "serving":
        serving_main(args)
        return
    if args.workload == "tune":
        tune_main(args)
        return
    if args.workload == "tt06":
        args.precision = args.precision or "mixed"
        tt06_main(args)
        return
    if args.workload == "tissue":
        args.precision = args.precision or "mixed"
        tissue_main(args)
        return
    args.precision = args.precision or "mixed"
    prec = cep.PREC_MIXED if args.precision == "mixed" else cep.PREC_FP64
    K = args.steps if args.steps is not None else 2
    W = args.warmup if args.warmup is not None else 1
    pop = make_population(args.ncell)
    n = len(pop["bcl"])
    nb_total = 200
    per_group = parse_group_cells(args.group_cells)
    b, pos = make_slot_batch(pop, prec, nb_total, per_group)
    log(f"bench: {n} cells in {b.n} slots, warmup {W} beat(s)")
    if W:
        b.run_beats(W)
    states_w = b.states()[pos]          # the state the timed beats start from (CPU baseline)
    b.kernel_time()                     # reset the event timer
    stats0 = b.stats()[pos]
    log(f"bench: timing {K} beat(s)")
    t0 = time.perf_counter()
    for _ in range(K):
        b.run_beats(1)
    t1 = time.perf_counter()
    nl, kavg_ms, ktot_ms = b.kernel_time()
    stats1 = b.stats()[pos]
    st = b.status()[pos]
    bm_timed = b.biomarker_array()[pos]
    nslots = b.n
    elapsed = t1 - t0
    value = n * K / elapsed
    flops_launch = flops_of(stats1 - stats0) / max(nl, 1)
    achieved = flops_launch / (kavg_ms * 1e-3) if kavg_ms > 0 else 0.0
    peak = PEAK_FP32 if prec == cep.PREC_MIXED else PEAK_FP64
    nrhs = float(np.sum((stats1 - stats0)[:, [1, 7]]))
    b.close()

    # the other precision on the same 10,000 cells: what mixed precision buys at full size
    other = {}
    if args.fp64_beats > 0:
        oprec = cep.PREC_FP64 if prec == cep.PREC_MIXED else cep.PREC_MIXED
        log(f"bench: {'FP64' if oprec == cep.PREC_FP64 else 'mixed'} batch of {n} cells, "
            f"1 + {args.fp64_beats} beat(s)")
        bo, _ = make_slot_batch(pop, oprec, nb_total, per_group)
        bo.run_beats(1)
        t2 = time.perf_counter()
        bo.run_beats(args.fp64_beats)
        t3 = time.perf_counter()
        bo.close()
        other = {"precision": "fp64" if oprec == cep.PREC_FP64 else "mixed",
                 "cell_beats_per_s": n * args.fp64_beats / (t3 - t2),
                 "beats_timed": args.fp64_beats, "after_beats": 1}

    # accuracy of the timed solver against the FP64 CPU oracle, BCL-stratified, at the beat
    # count the timed batch reached (or, with --check-beats, at that count)
    chk = {}
    if args.check_per_bcl > 0:
        cb = args.check_beats if args.check_beats > 0 else W + K
        chk = accuracy_check(pop, bm_timed if cb == W + K and not args.check_window else None,
                             args.check_per_bcl, cb, args.check_window or None, args.check_dump)
        plain = chk["fp64_plain_job_cell_beats_per_s"]
        chk["batched_mixed_over_plain_fp64"] = value / plain if prec == cep.PREC_MIXED else None
        if other.get("precision") == "fp64":
            chk["batched_mixed_over_batched_fp64"] = value / other["cell_beats_per_s"]
    lat = {"p50": None, "p99": None}
    if args.serve_jobs > 0:
        log(f"bench: serving session, {args.serve_jobs} jobs")
        lat = serving_latency(args.serve_jobs, args.serve_rate,
                              tuple(int(x) for x in args.serve_beats.split(",")))
    legs = {}
    if args.legs:
        legs = compact_legs(args)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline")
        cpu = cpu_baseline(pop, states_w, args.cpu_seconds, W)
    traffic = read_traffic()
    out = {
        "metric": "ORd cell-beats/s on 1 GPU at APD90 err <=1 ms; p50/p99 job latency",
        "value": value,
        "unit": "cell-beats/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-rates/f64-state" if prec == cep.PREC_MIXED else "f64",
        "data": "synthetic (seeded drug/IC50/conductance population)",
        "config": {"workload": "ORd-CiPA batched population, 10,000 cells = 50 drugs x 4 conc x 50 "
                               "variants, BCL {500,1000,2000} ms, BDF rtol 1e-6/atol 1e-8, step = 1 beat",
                   "cells": n, "beats_per_job": nb_total,
                   "layout": {"cells_per_wavefront_by_bcl": {str(int(k)): v for k, v in per_group.items()},
                              "default_cells_per_wavefront": 16, "slots": nslots}},
        "accuracy_vs_oracle": chk,
        "other_precision_full_batch": other,
        "failed_cells": int(np.sum(st != 0)),
        "latency_ms": lat,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic,
                     "traffic_source": "profile-derived: committed rocprofv3 FETCH_SIZE/WRITE_SIZE "
                                       "summary of this workload's launch, "
                                       "profiles/ord_bdf_population_10k_traffic.txt",
                     "kernel": f"ord_bdf_kernel<{'float' if prec == cep.PREC_MIXED else 'double'}, 16>",
                     "kernel_avg_ms": kavg_ms, "flops_per_launch": flops_launch,
                     "rhs_evals_per_launch": nrhs / max(nl, 1),
                     "valu_counters": read_valu("bdf") if prec == cep.PREC_MIXED else None},
        "other_workloads": legs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def compact_legs(args):
    """Short runs of the other configurations, timed by the same driver run: one beat of the
    200,000-cell TT06 population (mixed), 100 Strang steps of the 512 x 512 tissue (FP32 rates,
    FP64 state and CG; its Vm against an FP64 run of the same steps), and the first 100 of the
    1000 beats of the reference's ORd 2011 quinidine run (FP64) with two beats of a
    100,000-cell population of the same model."""
    legs = {}
    q, s0 = ord11_inputs()
    keys = [(i, c) for c in (3237.0, 6474.0, 9711.0) for i in range(len(q))]
    nb = 100
    log(f"bench: ORd 2011 leg, {len(keys)} cells x {nb} beats + 100000 cells x 2 beats")
    b = ord11_batch(len(keys), np.array([c for _, c in keys]), np.vstack([q[i] for i, _ in keys]),
                    cep.PREC_FP64, 1000, s0)
    t0 = time.perf_counter()
    b.run_beats(nb)
    t1 = time.perf_counter()
    bad = int(np.sum(b.status() != 0))
    b.close()
    rng = np.random.default_rng(20251208)
    npop = 100000
    bp = ord11_batch(npop, rng.uniform(0.0, 4 * 9711.0, npop), q[rng.integers(0, len(q), npop)],
                     cep.PREC_FP64, 2, s0)
    bp.kernel_time()
    t2 = time.perf_counter()
    bp.run_beats(2)
    t3 = time.perf_counter()
    _, pavg, _ = bp.kernel_time()
    bad += int(np.sum(bp.status() != 0))
    bp.close()
    legs["ord2011"] = {"workload": f"ORd 2011 M cell, quinidine IC50_samples10 x {{3237, 6474, 9711}} nM, BCL 2000, "
                                   f"FP64, the first {nb} of the reference run's 1000 beats",
                       "cell_beats_per_s": len(keys) * nb / (t1 - t0),
                       "reference_cell_beats_per_s": ORD11_REF_CELL_BEATS_PER_S,
                       "population_cell_beats_per_s": npop * 2 / (t3 - t2), "population_kernel_avg_ms": pavg,
                       "failed_cells": bad}
    pop = cep.workload.tt_population(20251205, args.tt_cells)
    herg = np.concatenate([pop["ic50"], pop["hill"]], axis=1)
    n = args.tt_cells
    log(f"bench: TT06 leg, {n} cells x 1 beat")
    b = cep.CellBatch(cep.MODEL_TT06, n, cep.PREC_MIXED)
    b.set_cells(celltype=pop["celltype"], conc=pop["conc"], herg=herg, gscale=pop["gscale"], bcl=pop["bcl"])
    b.set_run_config(dt=0.02, nbeats_total=500, nwin=500)
    b.init_consts()
    b.run_beats(1)
    b.kernel_time()
    t0 = time.perf_counter()
    b.run_beats(1)
    t1 = time.perf_counter()
    _, kavg, _ = b.kernel_time()
    bad = int(np.sum(b.status() != 0))
    b.close()
    fl = n * int(round(1000.0 / 0.02)) * F_TT_STEP
    legs["tt06"] = {"workload": f"TT06 {n} cells, Rush-Larsen dt 0.02 ms, mixed, beat 2",
                    "cell_beats_per_s": n / (t1 - t0), "kernel_avg_ms": kavg, "failed_cells": bad,
                    "roofline": {"bound": "valu", "achieved": fl / (kavg * 1e-3) / 1e12, "peak": PEAK_FP32 / 1e12,
                                 "unit": "TFLOP/s", "frac": fl / (kavg * 1e-3) / PEAK_FP32,
                                 "valu_counters": read_valu("tt")}}
    log("bench: tissue leg, 512 x 512, 20 + 100 steps")
    tis, _, _ = tissue_setup(TISSUE_NX, cep.PREC_MIXED)
    tis.step(20, TISSUE_DT, cep.SPLIT_STRANG)
    tis.kernel_times()
    t0 = time.perf_counter()
    iters = tis.step(100, TISSUE_DT, cep.SPLIT_STRANG)
    t1 = time.perf_counter()
    avg, _ = tis.kernel_times()
    nn = tis.nn
    vm = tis.vm()
    tis.close()
    t64, _, _ = tissue_setup(TISSUE_NX, cep.PREC_FP64)
    t64.step(120, TISSUE_DT, cep.SPLIT_STRANG)
    dv = np.abs(vm - t64.vm())
    t64.close()
    legs["tissue"] = {"workload": "512x512 Q1 monodomain, ORd-CiPA nodes, Strang dt 0.01 ms, FP32 rates, "
                                  "FP64 state and CG",
                      "node_steps_per_s": nn * 100 / (t1 - t0), "cg_iters_per_step": iters / 100,
                      "vm_linf_mv_vs_fp64": float(dv.max()),
                      "cg_spmv_gbs": nn * B_SPMV / (avg[1] * 1e-3) / 1e9 if avg[1] > 0 else None,
                      "kernels_avg_ms": {"tissue_react": float(avg[0]), "cg_spmv": float(avg[1]),
                                         "cg_update": float(avg[2])}}
    return legs


if __name__ == "__main__":
    main()
