/*
 * oracle.h -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * This library is a plain-C restatement of the reference's algorithms for the
 * cardiac cell-model hot path.  It exists only to check the HIP product path:
 * it is linked by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and by nothing else.  The product (the package's C-ABI
 * library) never calls into it.
 *
 * Pinning (what checks each part against something other than itself):
 *   ord_*  ORd-CiPA v1 2017 init/drug block/RHS -> golden vectors produced by the
 *          compiled reference model (oracle/_ref, tests/golden/ord_*.bin).
 *   tt_*   TT06 endo/epi/M init/RHS/Rush-Larsen -> golden vectors + 1-beat RL
 *          trajectories from the compiled reference model.
 *   ord11_*  ORd 2011 (static Hill) RHS + RL -> golden vectors from the compiled
 *          reference model (CEPPatch_CiPA_Sim copy).
 *   bdf_*  CVODE-style BDF: the reference only ships CVODE as a binary archive, so
 *          CVODE parity is UNPINNED.  The restatement is pinned against textbook
 *          problems (closed-form linear systems, Robertson, order-of-accuracy) and
 *          against fine-step integrations of the same cell model.
 *   bm_*   biomarker state machine -> synthetic closed-form traces with known
 *          APD/CaD/dVdt answers.
 *   fem_*  Q1 element matrices -> closed forms; CG -> dense direct solve.
 */
#ifndef CEP_ORACLE_H
#define CEP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- ORd-CiPA v1 2017 (CEP5Cell/cellmodels/ohara_rudy_cipa_v1_2017.cpp) */
#define ORD_NS 52
#define ORD_NC 206
#define ORD_NA 200
void ord_init_consts(double *C, double *S, double celltype);
/* herg[0..13] = IC50/Hill pairs, herg[14..18] = Kmax, Ku, n, halfmax, Vhalf */
void ord_apply_drug(double *C, double *S, double conc, const double *herg);
void ord_rates(double t, const double *C, double *R, const double *S, double *A);

/* ---------------- TT06 (CEP4Cell/cellmodels/ten_tusscher_model_2006_IK1Ko_*_units.cpp) */
#define TT_NS 19
#define TT_NC 53
#define TT_NA 70
/* celltype: 0 endo, 1 epi, 2 M */
void tt_init_consts(double *C, double *S, int celltype);
void tt_rates(double t, const double *C, double *R, const double *S, double *A, int celltype);
void tt_rush_larsen(double dt, const double *C, const double *R, double *S, const double *A);

/* ---------------- ORd 2011 (DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp) */
#define ORD11_NS 41
#define ORD11_NC 144
#define ORD11_NA 198
/* celltype is stored and then reset to 0 by the model's own initialisation (endo constants) */
void ord11_init_consts(double *C, double *S, double celltype);
void ord11_apply_dutta(double *C);
/* ic50[14]: (IC50, Hill) of ICaL, IK1, IKs, INa, INaL, Ito, hERG (IC50 CSV column order) */
void ord11_apply_drug(double *C, double conc, const double *ic50);
void ord11_rates(double t, const double *C, double *R, const double *S, double *A);
/* tissue_stim < 0: single-cell periodic stimulus; 0 / 1: the tissue stimulus off / on */
void ord11_rates_ex(double t, const double *C, double *R, const double *S, double *A, int tissue_stim);
void ord11_rush_larsen(double dt, const double *C, const double *R, double *S, const double *A);
double ord11_time_step(double t, double time_point, double max_dt, const double *C, const double *R);

/* per-beat features of the DrugSimulation driver (drug_sim.cpp:306-420) */
#define ORD11_BM_QNET_AP 0
#define ORD11_BM_QNET4_AP 1
#define ORD11_BM_INAL_AUC_AP 2
#define ORD11_BM_ICAL_AUC_AP 3
#define ORD11_BM_QNET_CL 4
#define ORD11_BM_QNET4_CL 5
#define ORD11_BM_INAL_AUC_CL 6
#define ORD11_BM_ICAL_AUC_CL 7
#define ORD11_BM_VM_PEAK 8
#define ORD11_BM_VM_VALLEY 9
#define ORD11_BM_DVMDT_REPOL 10
#define ORD11_BM_T_PEAK 11
#define ORD11_BM_VM_REPOL30 12
#define ORD11_BM_VM_REPOL50 13
#define ORD11_BM_VM_REPOL90 14
#define ORD11_BM_STEEPEST_PACE 15
#define ORD11_NBM 16
typedef struct {
  double celltype, conc, bcl;
  double ic50[14];
  double time_point, max_dt;   /* 25 ms, 1.0 ms in the reference */
  int dutta, nbeats, nwin;     /* nwin = last_drug_check_pace */
} ord11_cell_cfg_t;
long oracle_ord11_run(const ord11_cell_cfg_t *cfg, double *states_io, int use_states_in, double *res);
double oracle_ord11_run_many(const ord11_cell_cfg_t *cfgs, int ncell, int nthreads, const double *states_in,
                             double *res, double *states_out);

/* ---------------- CVODE-style BDF */
typedef void (*bdf_rhs_fn)(double t, const double *y, double *ydot, void *user);

typedef struct {
  double rtol, atol, hmax, hmin;
  int maxord, max_steps;   /* max_steps per call of bdf_advance */
} bdf_opts_t;

typedef struct {
  long nst, nfe, nsetups, nje, nni, ncfn, netf;
} bdf_stats_t;

typedef struct bdf_mem bdf_mem_t;
bdf_mem_t *bdf_create(int n, bdf_rhs_fn f, void *user, const bdf_opts_t *o);
void bdf_free(bdf_mem_t *m);
int bdf_init(bdf_mem_t *m, double t0, const double *y0);
/* advance to tout in normal mode: steps past tout and interpolates y(tout). returns 0 ok, <0 error */
int bdf_advance(bdf_mem_t *m, double tout, double *yout);
void bdf_get_stats(const bdf_mem_t *m, bdf_stats_t *s);
double bdf_tn(const bdf_mem_t *m);
double bdf_h(const bdf_mem_t *m);
int bdf_q(const bdf_mem_t *m);

/* ---------------- Shampine-Gordon Adams PECE (libs/ode.cpp) */
typedef void (*adams_rhs_fn)(double t, const double *y, double *yp, void *user);
typedef struct adams_mem adams_mem_t;
adams_mem_t *adams_create(int n, adams_rhs_fn f, void *user);
void adams_free(adams_mem_t *m);
/* one call of the reference's ode(): returns and stores iflag (2 = reached tout) */
int adams_ode(adams_mem_t *m, double *y, double *t, double tout, double relerr, double abserr, int *iflag);
void adams_counts(const adams_mem_t *m, long *nsteps, long *nfe, int *order);
/* the reference's Adams loop on a test problem (mode 0 linear, 1 Van der Pol, 2 ORd 2011) */
long oracle_adams_run(int mode, const double *in, double *out, long out_cap, long *nsteps, long *nfe);

/* ---------------- biomarkers (CEPCell_NEO/modules/cipa_drug_bench.cpp:210-381) */
typedef struct {
  double qnet, inal_auc, ical_auc, dvmdt_repol, dvmdt_max, vm_peak, vm_valley, vm_dia,
         apd90, apd50, ca_peak, ca_valley, ca_dia, cad90, cad50;
  int steepest_pace;
} bm_result_t;

typedef struct bm_state bm_state_t;
/* nbeats = num_pace1, nwin = number of trailing beats scanned (reference: 250) */
bm_state_t *bm_create(double bcl, double dt, double stim_start, int nbeats, int nwin, int print_freq);
void bm_free(bm_state_t *b);
void bm_begin(bm_state_t *b, double v0, double ca0);
/* one output sample: iout already incremented, t = tcurr, dvdt = RATES[V],
   qsum = INaL+ICaL+Ito+IKr+IKs+IK1, inal, ical */
void bm_sample(bm_state_t *b, int iout, double t, double v, double ca, double dvdt,
               double qsum, double inal, double ical);
void bm_result(const bm_state_t *b, bm_result_t *out);

/* ---------------- biomarker rules of the drugs_herg, benchpatch and population_tn2006 drivers
   (driver_rules.c); record layout shared with the library's cep_get_driver_rules */
#define DR_DH_APD 0
#define DR_DH_T_DEP 1
#define DR_DH_T_REP 2
#define DR_DH_NREP 3
#define DR_DH_QNET 4
#define DR_DH_INAL_AUC 5
#define DR_DH_ICAL_AUC 6
#define DR_BP_QNET 7
#define DR_BP_MAX_DVMDT 8
#define DR_BP_VM_PEAK 9
#define DR_BP_VM_DIA 10
#define DR_BP_APD90 11
#define DR_BP_APD50 12
#define DR_BP_CA_PEAK 13
#define DR_BP_CA_DIA 14
#define DR_BP_CAD90 15
#define DR_BP_CAD50 16
#define DR_N 17
#define DR_TN_APD 0
#define DR_TN_MEAN_CAI 1
typedef struct {
  double bcl, dt, v_apd90, v_apd50, v_prev;
  int nbeats, imax, w_prev, w_curr, w_cadia;
  double dh_t_dep, dh_t_rep, dh_apd, dh_q0, dh_n0, dh_c0, dh_q1, dh_n1, dh_c1;
  int dh_nrep;
  double bp_t_dep, bp_t_rep, bp_apd90, bp_apd50, bp_q0, bp_q1, bp_ca_dia, bp_ca_peak, bp_ca_amp;
  double bp_t1_cad90, bp_cad90, bp_t1_cad50, bp_cad50, bp_max_dvmdt, bp_vm_peak, bp_vm_dia;
  int bp_ca90_inc, bp_cad90_found, bp_ca50_inc, bp_cad50_found;
} dr_state_t;
void dr_init(dr_state_t *s, double bcl, double dt, int nbeats, double v_apd90, double v_apd50);
void dr_begin(dr_state_t *s, double v0);
void dr_sample(dr_state_t *s, int k, double t, double v, double ca, const double *acc, double dvdt_next);
void dr_result(const dr_state_t *s, double *out);
double dr_qinward(double inal_drug, double ical_drug, double inal_ctl, double ical_ctl);
typedef struct {
  double t_last, v_apd90, v_prev, t_dep, t_rep, apd, sum_ca;
  int m;
} tn_state_t;
void tn_init(tn_state_t *s, double bcl, int nbeats, double v_apd90, double v0);
void tn_sample(tn_state_t *s, double tcurr, double v, double ca);
void tn_result(const tn_state_t *s, double *out);

/* ---------------- drivers */
typedef struct {
  double celltype, conc, bcl, stim_amp_scale, dt, rtol, atol;
  double herg[19];
  double gscale[7];   /* GNa, GNaL, PCa, GKr, GKs, GK1, Gto */
  int nbeats, nwin, maxord;
  double v_apd90, v_apd50;   /* thresholds of the driver rules */
  double t_first;            /* first output time (0: dt) */
} ord_cell_cfg_t;

/* Runs one ORd-CiPA cell through nbeats beats with the BDF, output every dt,
   biomarkers over the trailing nwin beats.  Writes final states (52), the result,
   optional last-beat Vm trace (vm_trace[nsamp], nsamp = bcl/dt) and stats. */
int oracle_ord_run(const ord_cell_cfg_t *cfg, double *states_io, int use_states_in,
                   bm_result_t *res, double *vm_trace, bdf_stats_t *st);
/* the same run, also evaluating the driver rules (rules: DR_N doubles) */
int oracle_ord_run_rules(const ord_cell_cfg_t *cfg, double *states_io, int use_states_in,
                         bm_result_t *res, double *vm_trace, bdf_stats_t *st, double *rules);

typedef struct {
  double bcl, dt, stim_start_unused;
  int celltype, nbeats, nwin;
  double conc;
  double ic50[3], hill[3];      /* IKr, ICaL, INa */
  double gscale[10];            /* GNa GK1 GKr GKs Gto GCaL GbNa GbCa GpCa GpK */
} tt_cell_cfg_t;
int oracle_tt_run(const tt_cell_cfg_t *cfg, double *states_io, int use_states_in,
                  bm_result_t *res, double *vm_trace);
/* the same run with population_tn2006's rules at threshold v_apd90 (rules: 2 doubles) */
int oracle_tt_run_rules(const tt_cell_cfg_t *cfg, double *states_io, int use_states_in,
                        bm_result_t *res, double *vm_trace, double v_apd90, double *rules);

/* threads: run ncell independent ORd cells round-robin over nthreads host threads */
double oracle_ord_run_many(const ord_cell_cfg_t *cfgs, int ncell, int nthreads,
                           bm_result_t *res, double *states_out);
double oracle_tt_run_many(const tt_cell_cfg_t *cfgs, int ncell, int nthreads,
                          bm_result_t *res, double *states_out);
/* ncell ORd cells, cell c started from row c of states_io (52 doubles), final states written back */
double oracle_ord_run_many_from(const ord_cell_cfg_t *cfgs, int ncell, int nthreads, double *states_io,
                                bm_result_t *res);
/* ord_rates over n cells (rows of C and S), one thread; returns wall seconds */
double oracle_ord_rates_many(int n, const double *C, const double *S, double *R, double t);

/* ---------------- tissue (fem.cpp / heart.cpp) */
void fem_hex8_element(const double *X, const double *D, double *Me, double *Ke);
void fem_tet4_element(const double *X, const double *D, double *Me, double *Ke);
void fem_q4_element(const double *xe, const double *ye, double sxx, double sxy, double syy, double *Me,
                    double *Ke);
void fem_q1_element(double hx, double hy, double sxx, double sxy, double syy,
                    double *Me /*16*/, double *Ke /*16*/);
int cg_solve_csr(int n, const int *rowptr, const int *col, const double *val,
                 const double *b, double *x, double rtol, int maxit, int jacobi);

#ifdef __cplusplus
}
#endif
#endif
