/*
 * driver_rules.c -- ORACLE (test infrastructure only): the per-output biomarker rules of three
 * reference drivers other than CEPCell_NEO's, restated per output sample.
 *
 * drugs_herg (CEP7Cell/modules/drugs_herg.cpp:140-315), after each successful CVode call
 * (iout already incremented; the call that reaches iout == imax breaks out before the checks,
 * :221-226):
 *   - APD at the fixed threshold v_apd90: a downward crossing (v_prev > v_apd90 >= V) sets
 *     t_repolarize = t and apd = t_repolarize - t_depolarize; then an upward crossing
 *     (v_prev <= v_apd90 < V) sets t_depolarize = t (:247-255);
 *   - qnet, INaL and ICaL AUC as differences of the accumulator STATES between the outputs
 *     iout == (int)(BCL/dt)*(n-2) + (int)BCL and iout == (int)(BCL/dt)*(n-1) + (int)BCL,
 *     qnet divided by 1000 (:281-303, :306);
 *   - qinward = ((INaL_d/INaL_c) + (ICaL_d/ICaL_c))/2 against the concentration-0 run (:309-315).
 * benchpatch (CEPPatch_Population_Drugs/benchpatch.cpp:268-432): the output counter is tested
 * before its increment (:444), the first output is at t_out0 (0.01 in the reference, :297):
 *   - APD90 at v_apd90 on every output (:329-338); in the last cycle (t > BCL(n-1)) APD50 at
 *     v_apd50 with the same t_depolarize variable (:406-417), max RATES[V] evaluated at
 *     t_out = t + dt after the increment of t_out (:397-399), Vm peak (:402), both started from
 *     the initial V (:277-278);
 *   - Ca diastole at iout == (int)(BCL/dt)*(n-2) (:358-361), Ca peak over BCL(n-2) < t <
 *     BCL(n-1) (:363-370), CaD90 / CaD50 in the last cycle from the first output at or above
 *     dia + {0.1,0.5} amp to the first one below it afterwards (:373-394);
 *   - qnet window as drugs_herg with the pre-increment counter (:340-353); vm_dia = the final V.
 * population_tn2006 (CEP7Cell/modules/population_tn2006.cpp:143-191), after each Rush-Larsen step
 * with tcurr += dt: APD at v_apd90 as drugs_herg; mean Ca_i over the samples with
 * tcurr >= BCL(n-1), divided by (count - 1) as the reference does (:185-186).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

void dr_init(dr_state_t *s, double bcl, double dt, int nbeats, double v_apd90, double v_apd50)
{
  memset(s, 0, sizeof(*s));
  s->bcl = bcl; s->dt = dt; s->nbeats = nbeats; s->v_apd90 = v_apd90; s->v_apd50 = v_apd50;
  const int bs = (int)(bcl / dt);
  s->w_prev = bs * (nbeats - 2) + (int)bcl;
  s->w_curr = bs * (nbeats - 1) + (int)bcl;
  s->w_cadia = bs * (nbeats - 2);
  s->imax = (int)((nbeats * bcl) / dt);
}

void dr_begin(dr_state_t *s, double v0)
{
  s->v_prev = v0;
  s->bp_max_dvmdt = v0;      /* benchpatch.cpp:277 */
  s->bp_vm_peak = v0;        /* :278 */
}

/* one output: k = 1-based count of outputs so far (drugs_herg's iout), t = its time, v / ca the
   states there, acc[3] = STATES[qnet, INaL_AUC, ICaL_AUC], dvdt_next = RATES[V] evaluated at
   t + dt with the states at t (benchpatch's computeRates(t_out, ...)) */
void dr_sample(dr_state_t *s, int k, double t, double v, double ca, const double *acc, double dvdt_next)
{
  const double vp = s->v_prev;
  /* ---- drugs_herg: checks skipped on the output that ends the run */
  if (k < s->imax) {
    if (vp > s->v_apd90 && v <= s->v_apd90) {
      s->dh_t_rep = t;
      s->dh_apd = s->dh_t_rep - s->dh_t_dep;
      s->dh_nrep++;
    }
    if (vp <= s->v_apd90 && v > s->v_apd90) s->dh_t_dep = t;
    if (k == s->w_prev) { s->dh_q0 = acc[0]; s->dh_n0 = acc[1]; s->dh_c0 = acc[2]; }
    else if (k == s->w_curr) { s->dh_q1 = acc[0]; s->dh_n1 = acc[1]; s->dh_c1 = acc[2]; }
  }
  /* ---- benchpatch: counter before its increment */
  const int i = k - 1;
  const double n = s->nbeats, bcl = s->bcl;
  if (vp > s->v_apd90 && v <= s->v_apd90) {
    s->bp_t_rep = t;
    s->bp_apd90 = s->bp_t_rep - s->bp_t_dep;
  }
  if (vp <= s->v_apd90 && v > s->v_apd90) s->bp_t_dep = t;
  if (i == s->w_prev) { s->bp_q0 = acc[0]; }
  else if (i == s->w_curr) { s->bp_q1 = acc[0]; }
  if (i == s->w_cadia) s->bp_ca_dia = ca;
  if (t > bcl * (n - 2) && t < bcl * (n - 1)) {
    if (s->bp_ca_peak < ca) {
      s->bp_ca_peak = ca;
      s->bp_ca_amp = s->bp_ca_peak - s->bp_ca_dia;
    }
  }
  if (t > bcl * (n - 1)) {
    if (ca < (0.1 * s->bp_ca_amp) + s->bp_ca_dia && s->bp_ca90_inc && !s->bp_cad90_found) {
      s->bp_cad90 = t - s->bp_t1_cad90;
      s->bp_cad90_found = 1;
    } else if (ca >= (0.1 * s->bp_ca_amp) + s->bp_ca_dia && !s->bp_ca90_inc) {
      s->bp_t1_cad90 = t;
      s->bp_ca90_inc = 1;
    }
    if (ca < (0.5 * s->bp_ca_amp) + s->bp_ca_dia && s->bp_ca50_inc && !s->bp_cad50_found) {
      s->bp_cad50 = t - s->bp_t1_cad50;
      s->bp_cad50_found = 1;
    } else if (ca >= (0.5 * s->bp_ca_amp) + s->bp_ca_dia && !s->bp_ca50_inc) {
      s->bp_t1_cad50 = t;
      s->bp_ca50_inc = 1;
    }
    if (dvdt_next > s->bp_max_dvmdt) s->bp_max_dvmdt = dvdt_next;
    if (v > s->bp_vm_peak) s->bp_vm_peak = v;
    if (vp > s->v_apd50 && v <= s->v_apd50) {
      s->bp_t_rep = t;
      s->bp_apd50 = s->bp_t_rep - s->bp_t_dep;
    }
    if (vp <= s->v_apd50 && v > s->v_apd50) s->bp_t_dep = t;
  }
  s->bp_vm_dia = v;
  s->v_prev = v;
}

void dr_result(const dr_state_t *s, double *out)
{
  memset(out, 0, sizeof(double) * DR_N);
  out[DR_DH_APD] = s->dh_apd;
  out[DR_DH_T_DEP] = s->dh_t_dep;
  out[DR_DH_T_REP] = s->dh_t_rep;
  out[DR_DH_NREP] = s->dh_nrep;
  out[DR_DH_QNET] = (s->dh_q1 - s->dh_q0) / 1000.0;
  out[DR_DH_INAL_AUC] = s->dh_n1 - s->dh_n0;
  out[DR_DH_ICAL_AUC] = s->dh_c1 - s->dh_c0;
  out[DR_BP_QNET] = (s->bp_q1 - s->bp_q0) / 1000.0;
  out[DR_BP_MAX_DVMDT] = s->bp_max_dvmdt;
  out[DR_BP_VM_PEAK] = s->bp_vm_peak;
  out[DR_BP_VM_DIA] = s->bp_vm_dia;
  out[DR_BP_APD90] = s->bp_apd90;
  out[DR_BP_APD50] = s->bp_apd50;
  out[DR_BP_CA_PEAK] = s->bp_ca_peak;
  out[DR_BP_CA_DIA] = s->bp_ca_dia;
  out[DR_BP_CAD90] = s->bp_cad90;
  out[DR_BP_CAD50] = s->bp_cad50;
}

double dr_qinward(double inal_drug, double ical_drug, double inal_ctl, double ical_ctl)
{
  return ((inal_drug / inal_ctl) + (ical_drug / ical_ctl)) * 0.5;
}

/* ---- population_tn2006 */
void tn_init(tn_state_t *s, double bcl, int nbeats, double v_apd90, double v0)
{
  memset(s, 0, sizeof(*s));
  s->t_last = (nbeats - 1) * bcl;
  s->v_apd90 = v_apd90;
  s->v_prev = v0;
}

void tn_sample(tn_state_t *s, double tcurr, double v, double ca)
{
  if (s->v_prev > s->v_apd90 && v <= s->v_apd90) {
    s->t_rep = tcurr;
    s->apd = s->t_rep - s->t_dep;
  }
  if (s->v_prev <= s->v_apd90 && v > s->v_apd90) s->t_dep = tcurr;
  if (tcurr >= s->t_last) {
    s->sum_ca += ca;
    s->m++;
  }
  s->v_prev = v;
}

void tn_result(const tn_state_t *s, double *out)
{
  out[DR_TN_APD] = s->apd;
  out[DR_TN_MEAN_CAI] = s->sum_ca / (double)(s->m - 1);
}
