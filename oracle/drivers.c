/*
 * drivers.c -- ORACLE (test infrastructure only): per-cell drivers and the CPU
 * baseline (independent cells round-robined over host threads, the analogue of the
 * reference's MPI sample distribution, DrugSimulation/main.cpp:206-208).
 *
 * oracle_ord_run follows the CiPA drug loop: initConsts(celltype, conc, herg)
 * (CEP7Cell/modules/drugs_herg.cpp:147), BCL into the stimulus period, stimulus
 * amplitude scale (:148-150), CVode to every dt output point with max step dt
 * (:151-158,221-229), computeRates at the output time and biomarkers per sample
 * (CEPCell_NEO/modules/cipa_drug_bench.cpp:210-336).
 * oracle_tt_run follows the fixed-dt Rush-Larsen loop of
 * CEP7Cell/modules/population_tn2006.cpp:143-186 (computeRates(tcurr) then
 * solveAnalytical(dt), tcurr += dt), feeding the same biomarker state machine.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

typedef struct { double C[ORD_NC]; double A[ORD_NA]; } ord_ctx_t;

static void ord_f(double t, const double *y, double *ydot, void *u)
{
  ord_ctx_t *c = (ord_ctx_t *)u;
  ord_rates(t, c->C, ydot, y, c->A);
}

/* conductance scale slots for the population variants: GNa, GNaL, PCa, GKr, GKs, GK1, Gto */
static const int ord_gslot[7] = { 45, 153, 156, 159, 160, 161, 154 };

int oracle_ord_run(const ord_cell_cfg_t *cfg, double *states_io, int use_states_in,
                   bm_result_t *res, double *vm_trace, bdf_stats_t *st)
{
  return oracle_ord_run_rules(cfg, states_io, use_states_in, res, vm_trace, st, NULL);
}

/* stimulus of the single-cell drivers (ohara_rudy_cipa_v1_2017.cpp:950-951), for the RATES[V]
   that benchpatch evaluates one output step ahead */
static double ord_istim(const double *C, double t)
{
  const double start = C[12], end = C[13], amp = C[14], period = C[15], dur = C[16];
  return (t >= start && t <= end && (t - start) - floor((t - start) / period) * period <= dur) ? amp : 0.0;
}

int oracle_ord_run_rules(const ord_cell_cfg_t *cfg, double *states_io, int use_states_in,
                         bm_result_t *res, double *vm_trace, bdf_stats_t *st, double *rules)
{
  ord_ctx_t ctx;
  double S[ORD_NS], R[ORD_NS];
  ord_init_consts(ctx.C, S, cfg->celltype);
  ord_apply_drug(ctx.C, S, cfg->conc, cfg->herg);
  for (int k = 0; k < 7; ++k) ctx.C[ord_gslot[k]] *= (cfg->gscale[k] > 0 ? cfg->gscale[k] : 1.0);
  ctx.C[15] = cfg->bcl;
  ctx.C[14] *= cfg->stim_amp_scale;
  if (use_states_in) { memcpy(S, states_io, sizeof(S)); S[43] = cfg->conc; }

  bdf_opts_t o = { cfg->rtol, cfg->atol, cfg->dt, 0.0, cfg->maxord, 500 };
  bdf_mem_t *m = bdf_create(ORD_NS, ord_f, &ctx, &o);
  const double t0 = 0.0;
  bdf_init(m, t0, S);
  const int imax = (int)((cfg->nbeats * ctx.C[15]) / cfg->dt);
  const int beat_steps = (int)(ctx.C[15] / cfg->dt);
  /* Ca trace sampled every dt_write = 1 ms (reference default, CEP7Cell/commons/helper.cpp:13) */
  bm_state_t *b = bm_create(ctx.C[15], cfg->dt, ctx.C[12], cfg->nbeats, cfg->nwin, (int)((1. / cfg->dt) * 1.0));
  bm_begin(b, S[0], S[9]);
  dr_state_t dr;
  dr_init(&dr, ctx.C[15], cfg->dt, cfg->nbeats, cfg->v_apd90, cfg->v_apd50);
  dr_begin(&dr, S[0]);
  double tnext = cfg->t_first > 0.0 ? cfg->t_first : cfg->dt;
  int iout = 0, rc = 0;
  while (iout < imax) {
    rc = bdf_advance(m, tnext, S);
    if (rc != 0) break;
    ord_rates(tnext, ctx.C, R, S, ctx.A);
    const double tcurr = tnext;
    iout++;
    tnext += cfg->dt;
    const double *A = ctx.A;
    bm_sample(b, iout, tcurr, S[0], S[9], R[0],
              A[60] + A[83] + A[66] + A[90] + A[96] + A[98], A[60], A[83]);
    if (rules)
      dr_sample(&dr, iout, tcurr, S[0], S[9], &S[49], R[0] + ord_istim(ctx.C, tcurr) - ord_istim(ctx.C, tnext));
    if (vm_trace && iout > (cfg->nbeats - 1) * beat_steps) vm_trace[iout - 1 - (cfg->nbeats - 1) * beat_steps] = S[0];
  }
  if (res) bm_result(b, res);
  if (rules) dr_result(&dr, rules);
  if (st) bdf_get_stats(m, st);
  memcpy(states_io, S, sizeof(S));
  bm_free(b);
  bdf_free(m);
  return rc;
}

/* TT06 conductance slots: GNa GK1 GKr GKs Gto GCaL GbNa GbCa GpCa GpK */
static const int tt_gslot[10] = { 16, 13, 14, 15, 20, 18, 17, 19, 30, 32 };

int oracle_tt_run(const tt_cell_cfg_t *cfg, double *states_io, int use_states_in,
                  bm_result_t *res, double *vm_trace)
{
  return oracle_tt_run_rules(cfg, states_io, use_states_in, res, vm_trace, -40.0, NULL);
}

int oracle_tt_run_rules(const tt_cell_cfg_t *cfg, double *states_io, int use_states_in,
                        bm_result_t *res, double *vm_trace, double v_apd90, double *rules)
{
  double C[TT_NC], S[TT_NS], R[TT_NS], A[TT_NA];
  tt_init_consts(C, S, cfg->celltype);
  for (int k = 0; k < 10; ++k) C[tt_gslot[k]] *= (cfg->gscale[k] > 0 ? cfg->gscale[k] : 1.0);
  /* drug block on IKr, ICaL, INa by the Hill factor used for the CiPA model */
  const int dslot[3] = { 14, 18, 16 };
  for (int k = 0; k < 3; ++k)
    if (cfg->ic50[k] > 10E-14 && cfg->hill[k] > 10E-14)
      C[dslot[k]] *= pow(1. + pow(cfg->conc / cfg->ic50[k], cfg->hill[k]), -1);
  C[6] = cfg->bcl;
  if (use_states_in) memcpy(S, states_io, sizeof(S));
  memset(A, 0, sizeof(A));
  const int imax = (int)((cfg->nbeats * C[6]) / cfg->dt);
  const int beat_steps = (int)(C[6] / cfg->dt);
  bm_state_t *b = bm_create(C[6], cfg->dt, C[5], cfg->nbeats, cfg->nwin, (int)((1. / cfg->dt) * 1.0));
  bm_begin(b, S[0], S[3]);
  tn_state_t tn;
  tn_init(&tn, C[6], cfg->nbeats, v_apd90, S[0]);
  double tcurr = 0.0;
  int iout = 0;
  while (iout < imax) {
    tt_rates(tcurr, C, R, S, A, cfg->celltype);
    tt_rush_larsen(cfg->dt, C, R, S, A);
    tcurr += cfg->dt;
    iout++;
    if (rules) tn_sample(&tn, tcurr, S[0], S[3]);
    bm_sample(b, iout, tcurr, S[0], S[3], R[0], A[52] + A[54] + A[48] + A[49] + A[47], 0.0, A[52]);
    if (vm_trace && iout > (cfg->nbeats - 1) * beat_steps) vm_trace[iout - 1 - (cfg->nbeats - 1) * beat_steps] = S[0];
  }
  if (res) bm_result(b, res);
  if (rules) tn_result(&tn, rules);
  memcpy(states_io, S, sizeof(S));
  bm_free(b);
  return 0;
}

/* ---------------- thread pool baseline ---------------- */
typedef struct {
  int kind, ncell, nthreads, tid;
  const void *cfgs;
  bm_result_t *res;
  double *states_out;   /* kind 2: per-cell start states in, final states out */
} job_t;

static void *worker(void *p)
{
  job_t *j = (job_t *)p;
  for (int c = j->tid; c < j->ncell; c += j->nthreads) {
    if (j->kind == 0) {
      double S[ORD_NS];
      oracle_ord_run(&((const ord_cell_cfg_t *)j->cfgs)[c], S, 0, j->res ? &j->res[c] : NULL, NULL, NULL);
      if (j->states_out) memcpy(j->states_out + (size_t)c * ORD_NS, S, sizeof(S));
    } else if (j->kind == 2) {
      oracle_ord_run(&((const ord_cell_cfg_t *)j->cfgs)[c], j->states_out + (size_t)c * ORD_NS, 1,
                     j->res ? &j->res[c] : NULL, NULL, NULL);
    } else {
      double S[TT_NS];
      oracle_tt_run(&((const tt_cell_cfg_t *)j->cfgs)[c], S, 0, j->res ? &j->res[c] : NULL, NULL);
      if (j->states_out) memcpy(j->states_out + (size_t)c * TT_NS, S, sizeof(S));
    }
  }
  return NULL;
}

static double run_many(int kind, const void *cfgs, int ncell, int nthreads, bm_result_t *res, double *states_out)
{
  if (nthreads < 1) nthreads = 1;
  pthread_t th[256];
  job_t jobs[256];
  if (nthreads > 256) nthreads = 256;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int t = 0; t < nthreads; ++t) {
    jobs[t] = (job_t){ kind, ncell, nthreads, t, cfgs, res, states_out };
    pthread_create(&th[t], NULL, worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &b);
  return (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
}

double oracle_ord_run_many(const ord_cell_cfg_t *cfgs, int ncell, int nthreads, bm_result_t *res, double *states_out)
{
  return run_many(0, cfgs, ncell, nthreads, res, states_out);
}

/* as oracle_ord_run_many, each cell started from its row of states_io (the reference's restart
   from a saved state, cipa_drug_bench.cpp:151); final states are written back to the rows */
double oracle_ord_run_many_from(const ord_cell_cfg_t *cfgs, int ncell, int nthreads, double *states_io,
                                bm_result_t *res)
{
  return run_many(2, cfgs, ncell, nthreads, res, states_io);
}

double oracle_tt_run_many(const tt_cell_cfg_t *cfgs, int ncell, int nthreads, bm_result_t *res, double *states_out)
{
  return run_many(1, cfgs, ncell, nthreads, res, states_out);
}

/* Host timing of the reaction's right-hand sides: ord_rates for n cells (rows of C and S,
   row strides ORD_NC and ORD_NS), one thread; returns the wall seconds. */
double oracle_ord_rates_many(int n, const double *C, const double *S, double *R, double t)
{
  double A[ORD_NA];
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int i = 0; i < n; ++i)
    ord_rates(t, C + (size_t)i * ORD_NC, R + (size_t)i * ORD_NS, S + (size_t)i * ORD_NS, A);
  clock_gettime(CLOCK_MONOTONIC, &b);
  return (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
}
