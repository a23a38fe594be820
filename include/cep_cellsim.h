/*
 * cep_cellsim.h -- C ABI of the MI355X cardiac cell-population engine (libcepgpu.so).
 *
 * Drop-in boundary for the reference's CellML cell-model interface:
 *   class CellML (CEP5Cell/cellmodels/cellml.hpp:7-32): initConsts(type, conc, hill, drug),
 *   computeRates(TIME, CONSTANTS, RATES, STATES, ALGEBRAIC), solveAnalytical(dt), public
 *   STATES/RATES/ALGEBRAIC/CONSTANTS arrays;
 * driven by the CVODE loops of the drug/population drivers
 *   (CEP7Cell/modules/drugs_herg.cpp:140-300, CEPCell_NEO/modules/cipa_drug_bench.cpp:134-336,
 *   CEP7Cell/modules/population_tn2006.cpp:70-204).
 * Where the reference holds one model object per cell and integrates it on the CPU, this
 * ABI holds a batch of N cells resident in device memory and integrates all of them in
 * one launch.  Arrays crossing the ABI are host arrays in the reference's per-cell layout
 * (cell-major, STATES order of the reference enums); no device pointers or framework types
 * appear in the signatures.
 *
 * Errors: every call returns 0 on success, a negative CEP_E* code on failure;
 * cep_last_error() gives the message.  A cell whose integration fails is flagged in its
 * status word (CEP_ST_*) and stops; the other cells of the batch continue.
 * Threading: a batch is used by one host thread at a time; different batches may run
 * concurrently (each owns a HIP stream).
 */
#ifndef CEP_CELLSIM_H
#define CEP_CELLSIM_H

#ifdef __cplusplus
extern "C" {
#endif

#define CEP_MODEL_ORD_CIPA 1   /* O'Hara-Rudy CiPA v1 2017, 52 states (ohara_rudy_cipa_v1_2017.cpp) */
#define CEP_MODEL_TT06 2       /* ten Tusscher 2006 IK1Ko endo/epi/M, 19 states */
#define CEP_MODEL_ORD2011 3    /* O'Hara-Rudy 2011 of the DrugSimulation driver, 41 states
                                  (DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp), adaptive
                                  Rush-Larsen (solveAnalytical + set_time_step, :991-1137) */

#define CEP_PREC_FP64 0        /* all arithmetic in FP64 */
#define CEP_PREC_MIXED 1       /* gating/currents in FP32; state, Newton, LU, error control in FP64 */

#define CEP_E_ARG -1
#define CEP_E_HIP -2
#define CEP_E_STATE -3

#define CEP_ST_OK 0
#define CEP_ST_CONV_FAIL 1     /* Newton failed MAXNCF times (CVODE CV_CONV_FAILURE) */
#define CEP_ST_ERR_FAIL 2      /* error test failed MAXNEF times (CV_ERR_FAILURE) */
#define CEP_ST_NONFINITE 3
#define CEP_ST_MAX_STEPS 4     /* more than 500 steps between two output points (CV_TOO_MUCH_WORK) */
#define CEP_ST_IDLE 5          /* parked by cep_park_cells: no job in the cell, runs skip it */

/* biomarker record per cell (doubles), cipa_t of CEPCell_NEO/modules/commons.hpp:18-36 */
#define CEP_BM_QNET 0          /* integral of INaL+ICaL+Ito+IKr+IKs+IK1 over the selected beat (uA/uF ms) */
#define CEP_BM_INAL_AUC 1
#define CEP_BM_ICAL_AUC 2
#define CEP_BM_DVMDT_REPOL 3
#define CEP_BM_DVMDT_MAX 4
#define CEP_BM_VM_PEAK 5
#define CEP_BM_VM_VALLEY 6
#define CEP_BM_VM_DIA 7
#define CEP_BM_APD90 8
#define CEP_BM_APD50 9
#define CEP_BM_CA_PEAK 10
#define CEP_BM_CA_VALLEY 11
#define CEP_BM_CA_DIA 12
#define CEP_BM_CAD90 13
#define CEP_BM_CAD50 14
#define CEP_BM_STEEPEST_PACE 15
#define CEP_NBM 16

/* ORd 2011 record per cell: the DrugSimulation driver's features of the selected beat
   (drug_sim.cpp:306-420; the beat with the largest dVm/dt in [repol90, repol30] among the
   trailing nwin beats).  qnet*_ap accumulate while Vm > repol90, *_cl over the whole beat
   (C/F); an EAD is reported by the reference when dvmdt_repol > 0 (:424). */
#define CEP_BM11_QNET_AP 0
#define CEP_BM11_QNET4_AP 1
#define CEP_BM11_INAL_AUC_AP 2
#define CEP_BM11_ICAL_AUC_AP 3
#define CEP_BM11_QNET_CL 4
#define CEP_BM11_QNET4_CL 5
#define CEP_BM11_INAL_AUC_CL 6
#define CEP_BM11_ICAL_AUC_CL 7
#define CEP_BM11_VM_PEAK 8
#define CEP_BM11_VM_VALLEY 9
#define CEP_BM11_DVMDT_REPOL 10
#define CEP_BM11_T_PEAK 11
#define CEP_BM11_VM_REPOL30 12
#define CEP_BM11_VM_REPOL50 13
#define CEP_BM11_VM_REPOL90 14
#define CEP_BM11_STEEPEST_PACE 15

typedef struct cep_batch cep_batch;

/* Create a batch of ncells cells of one model.  Replaces `new <model>()` per cell
   (e.g. ohara_rudy_cipa_v1_2017.cpp:525-534). */
cep_batch *cep_batch_create(int model, int ncells, int precision);
void cep_batch_destroy(cep_batch *b);
/* Device memory of destroyed batches is kept by the library and reused by later batches
   (hipFree would wait for every stream on the device, stalling concurrent jobs).  This returns
   it to the device; it waits for the device like hipFree.  No counterpart in the reference
   (its cells live in host memory). */
int cep_release_cached_memory(void);
const char *cep_last_error(const cep_batch *b);
/* model, ncells, states per cell, constants per cell (CellML layout), algebraic slots */
int cep_batch_info(const cep_batch *b, int *model, int *ncells, int *nstates, int *nconsts, int *nalg);

/* Per-cell inputs for cells [first, first+count).  celltype: 0 endo, 1 epi, 2 M.
   ORd-CiPA: herg row = 14 IC50/Hill values in the IC50 CSV column order + Kmax, Ku, n,
   halfmax, Vhalf (initConsts(celltype, conc, herg, drug), ohara_rudy_cipa_v1_2017.cpp:819-845);
   gscale row = 7 population scales of GNa, GNaL, PCa, GKr, GKs, GK1, Gto
   (CEPPatch_Population_Drugs/benchpatch.cpp:211-241 style CONSTANTS[g] *= scale).
   TT06: herg row = IC50[IKr, ICaL, INa], Hill[IKr, ICaL, INa]; gscale row = 10 scales of
   GNa, GK1, GKr, GKs, Gto, GCaL, GbNa, GbCa, GpCa, GpK (population_tn2006.cpp:91-100).
   ORd 2011: herg row = the 14 IC50/Hill values of the IC50 CSV row (ICaL, IK1, IKs, INa, INaL,
   Ito, hERG; initConsts(celltype, conc, ic50, is_dutta), Ohara_Rudy_2011.cpp:722-740); gscale
   must be null; celltype is accepted but, as in the reference, the model's initialisation
   resets it, so every cell has the endocardial constants.
   bcl: basic cycle length (ms) -> stimulus period.  Null gscale = all ones. */
int cep_set_cells(cep_batch *b, int first, int count, const double *celltype, const double *conc,
                  const double *herg, int herg_stride, const double *gscale, int gscale_stride,
                  const double *bcl);

/* Model options (no reference counterpart as a call; they are the driver parameters of
   DrugSimulation/modules/drug_sim.cpp:74-81 and the Is_Dutta deck entry):
   CEP_OPT_DUTTA (ORd 2011): nonzero applies ___applyDutta (Ohara_Rudy_2011.cpp:460-467) after
   the constants, as the reference does (it scales the CONSTANTS slots 98, 96, 100, 81, 64);
   CEP_OPT_TIME_POINT (ORd 2011): the fixed-0.005-ms window at the start of every beat (25 ms).
   Call before cep_init_consts. */
#define CEP_OPT_DUTTA 1
#define CEP_OPT_TIME_POINT 2
/* CEP_OPT_INTEGRATOR (ORd 2011): CEP_INTEGRATOR_RUSH_LARSEN (the DrugSimulation loop, default)
   or CEP_INTEGRATOR_ADAMS: the Shampine-Gordon Adams PECE solver of libs/ode.cpp driven as
   CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:119-146 does (ode() to every output point
   dt, relerr = rtol, abserr = atol of cep_set_run_config; no features, optional Vm trace). */
#define CEP_OPT_INTEGRATOR 3
#define CEP_INTEGRATOR_RUSH_LARSEN 0
#define CEP_INTEGRATOR_ADAMS 1
/* CEP_OPT_TIME_SERIES (ORd-CiPA): nonzero records the time series of the selected (steepest)
   beat and the per-pace dvmdt_repol history that the reference driver writes with
   Is_Print_Graph (CEPCell_NEO/modules/cipa_drug_bench.cpp:168-199,230-235,262,338-350):
   t, Vm, Ca_i and INaL ICaL Ito IKr IKs IK1 at every dt_write (1 ms) point of the beat. */
#define CEP_OPT_TIME_SERIES 4
/* CEP_OPT_ORD11_WAVES (ORd 2011, Rush-Larsen): wavefronts per 64 cells.  0 (default) chooses:
   four (Na / Ito gates, Ca / K gates, Na and K currents, Ca currents with the rates and
   features, on four SIMDs) while four times the wave count fits one wave per SIMD, else two
   (gate kinetics / currents) while that fits, else one; 1, 2 or 4 forces it.  Results are the
   same either way up to floating-point contraction.  No reference counterpart (a launch-shape
   choice). */
#define CEP_OPT_ORD11_WAVES 5
/* CEP_OPT_DRIVER_RULES (ORd-CiPA, TT06): nonzero also evaluates, at every output point, the
   biomarker rules of the reference's other drivers (read with cep_get_driver_rules).  Enable before
   cep_init_consts; nbeats_total of cep_set_run_config is the drivers' num_pace1 (>= 2).
   ORd-CiPA: CEP7Cell/modules/drugs_herg.cpp:221-315 (APD from crossings of a fixed v_apd90, qnet
   and AUCs as differences of STATES[qnet, INaL_AUC, ICaL_AUC] between outputs (BCL/dt)(n-2)+BCL
   and (BCL/dt)(n-1)+BCL) and CEPPatch_Population_Drugs/benchpatch.cpp:320-432 (crossing APD90 and
   APD50, Ca diastole, the previous beat's Ca peak, CaD90/CaD50 from threshold crossings, max dV/dt
   and Vm peak of the last cycle, the final V).  TT06: CEP7Cell/modules/population_tn2006.cpp:
   143-191 (crossing APD, mean Ca_i of the last beat).
   CEP_OPT_V_APD90 (ORd-CiPA, TT06), CEP_OPT_V_APD50 (ORd-CiPA): the thresholds (mV; default -40,
   the reference's default v_apd90, CEP7Cell/commons/helper.cpp:21).
   CEP_OPT_FIRST_OUTPUT (ORd-CiPA): time of the first output point (ms; 0 = dt); benchpatch starts
   its outputs at 0.01 (benchpatch.cpp:297). */
#define CEP_OPT_DRIVER_RULES 6
#define CEP_OPT_V_APD90 7
#define CEP_OPT_V_APD50 8
#define CEP_OPT_FIRST_OUTPUT 9
/* CEP_OPT_KT (ORd-CiPA): the drug trapping rate Kt of the Markov IKr block (1/ms), CONSTANTS[100]
   of ohara_rudy_cipa_v1_2017.cpp (0 in its ___initConsts, :690, and in every reference driver);
   applies to every cell of the batch; call before cep_init_consts. */
#define CEP_OPT_KT 10
int cep_set_model_option(cep_batch *b, int option, double value);

/* Integration / biomarker configuration shared by the batch.
   dt: output step (ms; also CVODE max step for ORd, the Rush-Larsen step for TT06, the
   largest adaptive step max_time_step for ORd 2011, 1.0 in the reference);
   rtol/atol: BDF scalar tolerances (CVodeSStolerances); maxord: BDF max order (5);
   nbeats_total: beats of the whole job (biomarkers scan its trailing nwin beats);
   stim_amp_scale: multiplies the stimulus amplitude (drugs_herg.cpp:150). */
int cep_set_run_config(cep_batch *b, double dt, double rtol, double atol, int maxord,
                       int nbeats_total, int nwin, double stim_amp_scale);

/* Build per-cell constants and initial states, reset integrators and biomarkers
   (initConsts + CVodeInit/CVodeReInit). */
int cep_init_consts(cep_batch *b);

/* CONSTANTS of one cell in the reference's CellML layout (206 ORd / 53 TT06). */
int cep_get_constants(const cep_batch *b, int cell, double *out);

/* computeRates for every cell at time t: states/rates n x nstates, algebraic n x nalg
   (reference ALGEBRAIC indices; may be null).  Evaluated on the device. */
int cep_compute_rates(cep_batch *b, double t, const double *states, double *rates, double *algebraic);

/* cep_compute_rates without any device allocation on the call: no counterpart in the reference.
   For callers that cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  The states / rates / ALGEBRAIC staging is a fixed
   device array of the library (4.98 MB on each GPU it runs on, held until the process ends), so a
   batch may have at most CEP_STATIC_RATES_MAX_CELLS cells (CEP_E_ARG beyond).  Every call on one
   GPU shares that array: calls that overlap on one GPU race.  Same results as cep_compute_rates. */
#define CEP_STATIC_RATES_MAX_CELLS 2048
int cep_compute_rates_static(cep_batch *b, double t, const double *states, double *rates, double *algebraic);

/* Advance every cell by nbeats beats of its own BCL (device-resident state, biomarkers
   updated on the device).  Blocks until done.  The _async form returns after the launch on
   the batch stream; cep_sync waits for it. */
int cep_run_beats(cep_batch *b, int nbeats);
/* Advance every cell by nout output points of dt.  ORd-CiPA: the BDF (a CVode call per output
   point, biomarkers sampled at each); ORd 2011 with the Adams integrator: the reference loop's
   ode() calls (CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:132-140).  Blocks. */
int cep_run_outputs(cep_batch *b, int nout);
/* ORd-CiPA: CVode(cvode_mem, tout, ...) in CV_NORMAL mode for every cell (CEP7Cell/modules/
   drugs_herg.cpp:221-229): integrate to the output point tout = t_first + k dt (t_first =
   CEP_OPT_FIRST_OUTPUT, default dt), sampling the biomarkers at every output point on the way.
   Cells already at or past tout are left as they are.  Blocks. */
int cep_integrate_bdf(cep_batch *b, double tout);
/* One fixed-dt step of every cell at time t: computeRates(t) then solveAnalytical(dt) (TT06,
   ten_tusscher_model_2006_IK1Ko_endo_units.cpp:272-404 as population_tn2006.cpp:145-155 calls
   them; ORd 2011, Ohara_Rudy_2011.cpp:742-1095 with the single-cell stimulus).  ORd-CiPA has an
   empty solveAnalytical (ohara_rudy_cipa_v1_2017.cpp:1113-1115); its step here is Rush-Larsen for
   the gates, backward Euler for the Markov IKr block and forward Euler for the rest.  No
   biomarkers; the states stay on the device (cep_get_states).  Blocks. */
int cep_step_rl(cep_batch *b, double t, double dt);
int cep_run_beats_async(cep_batch *b, int nbeats);
int cep_sync(cep_batch *b);
/* Non-blocking completion test of the last cep_run_beats_async: 1 while it runs, 0 when
   done (then equivalent to cep_sync), <0 on error.  No reference counterpart: the
   reference drivers run one cell to completion per process; the serving loop uses this to
   overlap jobs on the device. */
int cep_query(cep_batch *b);
/* Device bytes a batch of ncells cells takes (state, integrator workspace, biomarker
   traces for beats up to max_bcl at output step dt, optional Vm trace); used for
   memory-based admission of jobs.  No reference counterpart. */
long long cep_batch_device_bytes(int model, int ncells, double max_bcl, double dt, int vm_trace);

/* Cells restarted and parked one by one, so that one resident batch serves many jobs: a job
   takes free cells at a beat boundary and gives them back when it ends.  No reference
   counterpart (each reference driver builds its cells, runs them to the end and deletes
   them: `new ohara_rudy_cipa_v1_2017()` ... `initConsts(type, conc, herg, drug)` per cell,
   CEP7Cell/modules/drugs_herg.cpp:113-145); a restart is that sequence for the listed cells
   inside a live batch.
   cep_restart_cells: each of the m distinct cells idx[] takes the parameters last given by
   cep_set_cells, its model's initial states, a fresh integrator (CVodeInit at t = 0) and
   biomarker machine, status CEP_ST_OK, and its own biomarker window: the last nwin of
   nbeats_total beats counted from this restart (read by the beat runs, cep_run_beats and
   cep_run_beats_async; the output-point runs keep the batch's window).  Every other cell is left as it is.  ORd-CiPA
   (BDF) and TT06 batches, after cep_init_consts and between runs (CEP_E_STATE while a run
   is in flight); a cell's BCL may not exceed the longest BCL of the batch at cep_init_consts
   (the Ca trace is sized for it); not with the driver rules on (they count beats for the
   whole batch).  Blocks until the restart is on the device.
   cep_park_cells: status CEP_ST_IDLE for the listed cells; every run skips them until they
   are restarted. */
int cep_restart_cells(cep_batch *b, const int *idx, int m, int nbeats_total, int nwin);
int cep_park_cells(cep_batch *b, const int *idx, int m);

/* Current STATES (n x nstates); set_states also restarts the integrators at t = 0 from
   the given states (the reference's Is_Using_Output restart, cipa_drug_bench.cpp:151). */
int cep_get_states(const cep_batch *b, double *out);
int cep_set_states(cep_batch *b, const double *in);

int cep_get_biomarkers(const cep_batch *b, double *out);        /* n x CEP_NBM */
int cep_get_status(const cep_batch *b, int *status);             /* n */
/* n x 8: steps, rhs evals, jacobians, setups, newton iters, conv fails, err fails, outputs */
int cep_get_stats(const cep_batch *b, long long *out);
/* Vm at every output point of the most recent beat (enable before running). */
int cep_enable_vm_trace(cep_batch *b, int enable);
int cep_get_vm_trace(const cep_batch *b, double *out, int *nsamp);  /* n x nsamp */
/* The selected beat's time series of one cell (CEP_OPT_TIME_SERIES): *n samples (the first
   max_n are copied; ires is max_n x 6: INaL ICaL Ito IKr IKs IK1), *pace = the selected pace
   (may be null).  All output arrays null: size query.  The reference's vm_data / cai_data /
   ires_data of cipa_result (cipa_drug_bench.cpp:230-235,264-278). */
int cep_get_time_series(const cep_batch *b, int cell, int max_n, double *t, double *vm, double *cai,
                        double *ires, int *n, int *pace);
/* dvmdt_repol of every eligible pace of the window, in pace order (the reference's
   max_dvdt_repol file, cipa_drug_bench.cpp:262): *n entries, the first max_n copied. */
int cep_get_dvmdt_repol_history(const cep_batch *b, int cell, int max_n, int *paces, double *values, int *n);

/* driver-rule records (CEP_OPT_DRIVER_RULES on): ORd-CiPA n x CEP_NDR, TT06 n x CEP_NDR_TT06 */
#define CEP_DR_DH_APD 0        /* drugs_herg: apd of the last repolarisation through v_apd90 */
#define CEP_DR_DH_T_DEP 1      /*   its t_depolarize, t_repolarize, and the count of crossings */
#define CEP_DR_DH_T_REP 2
#define CEP_DR_DH_NREP 3
#define CEP_DR_DH_QNET 4       /*   (qnet_curr - qnet_prev) / 1000 (drugs_herg.cpp:306) */
#define CEP_DR_DH_INAL_AUC 5   /*   INaL_AUC and ICaL_AUC over the same window */
#define CEP_DR_DH_ICAL_AUC 6
#define CEP_DR_BP_QNET 7       /* benchpatch: qnet_tot, max_dvmdt, vm_peak, vm_dia, apd90, apd50, */
#define CEP_DR_BP_MAX_DVMDT 8  /*   ca_peak, ca_dia, cad90, cad50 (benchpatch.cpp:420-432) */
#define CEP_DR_BP_VM_PEAK 9
#define CEP_DR_BP_VM_DIA 10
#define CEP_DR_BP_APD90 11
#define CEP_DR_BP_APD50 12
#define CEP_DR_BP_CA_PEAK 13
#define CEP_DR_BP_CA_DIA 14
#define CEP_DR_BP_CAD90 15
#define CEP_DR_BP_CAD50 16
#define CEP_NDR 17
#define CEP_DR_TN_APD 0        /* population_tn2006: ap_duration, mean_ca_i */
#define CEP_DR_TN_MEAN_CAI 1
#define CEP_NDR_TT06 2
int cep_get_driver_rules(const cep_batch *b, double *out);
/* qinward of every cell against control[i], the index of its concentration-0 cell in the same
   batch (drugs_herg.cpp:309-315): ((INaL_d / INaL_c) + (ICaL_d / ICaL_c)) / 2 over the drugs_herg
   AUC windows, computed on the device; control[i] < 0 gives NaN.  ORd-CiPA, driver rules on. */
int cep_get_qinward(cep_batch *b, const int *control, double *out);

/* average duration (ms) of the integration kernel over the launches since the last
   call, timed with HIP events on the batch stream; returns launches counted */
int cep_kernel_time(cep_batch *b, double *avg_ms, double *total_ms);

/* ---- launch shapes and the autotuner.  No reference counterpart: how the library maps cells
   onto the GPU.  Every shape computes the same arithmetic; the ORd-CiPA group order and the
   TT06 / ORd 2011 register split give bitwise the same results, the ORd 2011 wavefront split
   the same up to floating-point contraction.  Value 0 is automatic everywhere. */
#define CEP_SHAPE_WAVES_PER_EU 0   /* TT06, ORd 2011 one-wave kernel: 1 (all 512 registers, no
                                      spills) or 2 (256 registers, spills to scratch, a second
                                      wave per SIMD hides latency); automatic 2.  ORd-CiPA: 1. */
#define CEP_SHAPE_CELLS_PER_WAVE 1 /* ORd 2011: 64, 32 or 16 (the step over 1, 2 or 4 wavefronts
                                      exchanging through LDS; CEP_OPT_ORD11_WAVES); ORd-CiPA: 16
                                      (16 integrator and 16 output lanes per wavefront); TT06: 64 */
#define CEP_SHAPE_BLOCK 2          /* TT06: threads per workgroup, 64, 128 or 256 (automatic 256) */
#define CEP_SHAPE_GROUP_ORDER 3    /* ORd-CiPA whole-beat launches: CEP_ORDER_STORAGE or
                                      CEP_ORDER_LONGEST_FIRST (groups of 16 cells by their longest
                                      BCL); automatic: longest first when the groups outnumber the
                                      workgroups the device holds at once */
#define CEP_NSHAPE 4
#define CEP_ORDER_STORAGE 1
#define CEP_ORDER_LONGEST_FIRST 2
int cep_set_launch_shape(cep_batch *b, int knob, int value);
/* the value the next launch uses (automatic resolved) */
int cep_get_launch_shape(cep_batch *b, int knob, int *value);
/* Times every candidate shape of the batch's model on trial_beats beats of its own cells from its
   current state (HIP events on the batch stream), keeps the fastest, and restores the batch's
   device state bitwise, so the tuning run leaves no trace in the results.  report (may be null):
   per candidate CEP_NSHAPE shape values then its ms, (CEP_NSHAPE + 1) doubles, at most
   max_candidates rows; *ncandidates gets the count.  cache_path (may be null): the choice is
   stored there under the key (device architecture, CU count, model, precision, integrator, cell
   count rounded up to a power of two), replacing an older entry.  Allocates a copy of the
   batch's device memory for the duration of the call.  Returns the index of the chosen row. */
int cep_autotune(cep_batch *b, int trial_beats, const char *cache_path, double *report, int max_candidates,
                 int *ncandidates);
/* Applies the cached choice for the batch's key to every shape still automatic: 1 if an entry was
   found, 0 if not (or no file).  cep_init_consts does this itself when the environment variable
   CEP_TUNE_CACHE names a cache file. */
int cep_tune_lookup(cep_batch *b, const char *cache_path);

#ifdef __cplusplus
}
#endif
#endif
