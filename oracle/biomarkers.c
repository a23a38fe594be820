/*
 * biomarkers.c -- ORACLE (test infrastructure only): the 14-biomarker state machine
 * of CEPCell_NEO/modules/cipa_drug_bench.cpp:210-381, restated per output sample.
 *
 * Per output sample (after each successful CVode call to t = iout*dt):
 *   qnet/INaL/ICaL AUC by the rectangle rule (:225-227); Ca samples of the trailing
 *   window kept per beat (:230-235); beat-boundary processing at iout % (BCL/dt) == 0
 *   (:238-291): CaD50/CaD90 from the Ca trace split at the Ca peak, steepest-
 *   repolarisation beat selection, first-window-beat initialisation; within a beat
 *   (:296-334): max dV/dt, Vm peak sampled where iout % (int)((BCL*pace+stim_start+6)/dt)
 *   == 0 (kept literally, including its effect on pace 0), repol30/50/90 levels, dV/dt
 *   in [repol90, repol30], APD50/APD90 as the last time V sits in (level-2, level),
 *   Ca peak and the 50/90% Ca levels.
 * The reference runs this only for num_pace1 > 250 over the last 250 paces; here the
 * window length nwin is a parameter (nwin <= nbeats), 250 reproducing the reference.
 * CaD bounds persist across beats as in the reference (function-scope variables).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct bm_state {
  double bcl, dt, stim_start;
  int nbeats, nwin, print_freq, beat_steps, win_start;
  int pace, steepest_pace, eligible;
  double qnet, inal_auc, ical_auc;
  double t_depol, repol30, repol50, repol90, t_ca_peak, ca_amp50, ca_amp90;
  double cad50_prev, cad50_curr, cad90_prev, cad90_curr;
  bm_result_t temp, best;
  double *ca_t, *ca_v;
  int nca, cap;
};

static void res_init(bm_result_t *r, double v, double ca)
{
  r->qnet = 0.; r->inal_auc = 0.; r->ical_auc = 0.;
  r->dvmdt_repol = -999; r->dvmdt_max = -999; r->vm_peak = -999; r->vm_valley = v;
  r->vm_dia = -999; r->apd90 = 0.; r->apd50 = 0.; r->ca_peak = -999; r->ca_valley = ca;
  r->ca_dia = -999; r->cad90 = 0.; r->cad50 = 0.;
}

bm_state_t *bm_create(double bcl, double dt, double stim_start, int nbeats, int nwin, int print_freq)
{
  bm_state_t *b = (bm_state_t *)calloc(1, sizeof(bm_state_t));
  b->bcl = bcl; b->dt = dt; b->stim_start = stim_start; b->nbeats = nbeats;
  b->nwin = (nwin > nbeats) ? nbeats : nwin;
  b->print_freq = print_freq > 0 ? print_freq : 1;
  b->beat_steps = (int)(bcl / dt);
  b->win_start = nbeats - b->nwin;
  b->cap = b->beat_steps / b->print_freq + 4;
  b->ca_t = (double *)malloc(sizeof(double) * b->cap);
  b->ca_v = (double *)malloc(sizeof(double) * b->cap);
  return b;
}

void bm_free(bm_state_t *b)
{
  if (!b) return;
  free(b->ca_t); free(b->ca_v); free(b);
}

void bm_begin(bm_state_t *b, double v0, double ca0)
{
  res_init(&b->best, v0, ca0);
  res_init(&b->temp, v0, ca0);
  b->qnet = b->inal_auc = b->ical_auc = 0.;
  b->pace = 0; b->steepest_pace = 0; b->eligible = 0;
  b->t_depol = (b->bcl * b->pace) + b->stim_start;
  b->t_ca_peak = 0.; b->ca_amp50 = b->ca_amp90 = 0.;
  b->cad50_prev = b->cad50_curr = b->cad90_prev = b->cad90_curr = 0.;
  b->repol30 = b->repol50 = b->repol90 = 0.;
  b->nca = 0;
}

void bm_sample(bm_state_t *b, int iout, double t, double v, double ca, double dvdt,
               double qsum, double inal, double ical)
{
  b->qnet += qsum * b->dt;
  b->inal_auc += inal * b->dt;
  b->ical_auc += ical * b->dt;
  if (b->pace >= b->win_start && iout % b->print_freq == 0 && b->nca < b->cap) {
    b->ca_t[b->nca] = t; b->ca_v[b->nca] = ca; b->nca++;
  }
  if (iout > 0 && iout % b->beat_steps == 0) {
    if (b->eligible && b->pace >= b->win_start) {
      for (int k = 0; k < b->nca; ++k) {
        if (b->ca_t[k] < b->t_ca_peak) {
          if (b->ca_v[k] < b->ca_amp50) b->cad50_prev = b->ca_t[k];
          if (b->ca_v[k] < b->ca_amp90) b->cad90_prev = b->ca_t[k];
        } else {
          if (b->ca_v[k] > b->ca_amp50) b->cad50_curr = b->ca_t[k];
          if (b->ca_v[k] > b->ca_amp90) b->cad90_curr = b->ca_t[k];
        }
      }
      b->temp.cad50 = b->cad50_curr - b->cad50_prev;
      b->temp.cad90 = b->cad90_curr - b->cad90_prev;
      b->temp.qnet = b->qnet;
      b->temp.inal_auc = b->inal_auc;
      b->temp.ical_auc = b->ical_auc;
      b->temp.vm_dia = v;
      b->temp.ca_dia = ca;
      if (b->temp.dvmdt_repol > b->best.dvmdt_repol) {
        b->steepest_pace = b->pace;
        b->best = b->temp;
      }
    }
    b->pace++;
    if (b->pace == b->win_start + 1) {
      b->best = b->temp;
      b->steepest_pace = b->pace - 1;
    }
    b->qnet = b->inal_auc = b->ical_auc = 0.;
    if (b->pace >= b->win_start) {
      res_init(&b->temp, v, ca);
      b->nca = 0;
      b->t_ca_peak = t;
      b->t_depol = (b->bcl * b->pace) + b->stim_start;
      b->eligible = 0;
    }
  }
  if (b->pace >= b->win_start) {
    if (b->temp.dvmdt_max < dvdt) b->temp.dvmdt_max = dvdt;
    const int peak_idx = (int)(((b->bcl * b->pace) + (b->stim_start + 6.)) / b->dt);
    if (iout % peak_idx == 0) {
      b->temp.vm_peak = v;
      if (b->temp.vm_peak > 0.) {
        b->repol30 = b->temp.vm_peak - (0.3 * (b->temp.vm_peak - b->temp.vm_valley));
        b->repol50 = b->temp.vm_peak - (0.5 * (b->temp.vm_peak - b->temp.vm_valley));
        b->repol90 = b->temp.vm_peak - (0.9 * (b->temp.vm_peak - b->temp.vm_valley));
        b->eligible = 1;
      }
    } else if (b->eligible && t > (b->bcl * b->pace) + (b->stim_start + 6.)) {
      if (b->repol30 >= v && v >= b->repol90 && b->temp.dvmdt_repol < dvdt) b->temp.dvmdt_repol = dvdt;
      if (b->repol50 > v && v > b->repol50 - 2) b->temp.apd50 = t - b->t_depol;
      if (b->repol90 > v && v > b->repol90 - 2) b->temp.apd90 = t - b->t_depol;
      if (b->temp.ca_peak < ca) {
        b->temp.ca_peak = ca;
        b->ca_amp50 = b->temp.ca_peak - (0.5 * (b->temp.ca_peak - b->temp.ca_valley));
        b->ca_amp90 = b->temp.ca_peak - (0.9 * (b->temp.ca_peak - b->temp.ca_valley));
        b->t_ca_peak = t;
      }
    }
  }
}

void bm_result(const bm_state_t *b, bm_result_t *out)
{
  *out = b->best;
  out->steepest_pace = b->steepest_pace;
}
