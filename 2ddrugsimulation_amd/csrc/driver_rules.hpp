// driver_rules.hpp -- the biomarker rules of three reference drivers besides CEPCell_NEO's,
// evaluated per output point on the device.
//
// drugs_herg (CEP7Cell/modules/drugs_herg.cpp:221-315): APD from crossings of a fixed v_apd90
// (:247-255), qnet / INaL / ICaL AUC as differences of the accumulator STATES between outputs
// (BCL/dt)(n-2)+BCL and (BCL/dt)(n-1)+BCL (:281-303), no checks on the output that ends the
// run (:226), qinward against the concentration-0 run (:309-315).
// benchpatch (CEPPatch_Population_Drugs/benchpatch.cpp:320-432): the same crossings on every
// output, in the last cycle APD50 at v_apd50 through the same depolarisation time, max RATES[V]
// taken one output step ahead, Vm peak (both started from the initial V), Ca diastole, the
// previous beat's Ca peak and CaD90 / CaD50 from threshold crossings, the qnet window read with
// the counter before its increment, the final V.
// population_tn2006 (CEP7Cell/modules/population_tn2006.cpp:143-191): crossing APD and the mean
// Ca_i of the last beat divided by (samples - 1).
// State per cell: DR_SN doubles, SoA (slot k of cell c at d[k * pitch + c]).
#pragma once
#include "common.hpp"

namespace cep {

constexpr int DR_SN = 28;      // state doubles per cell
constexpr int DR_N = 17;       // record of the ORd-CiPA rules
constexpr int TN_N = 2;        // record of the TT06 population rules

struct DriverRules {
  double v_prev, dh_t_dep, dh_t_rep, dh_apd, dh_nrep, dh_q0, dh_n0, dh_c0, dh_q1, dh_n1, dh_c1;
  double bp_t_dep, bp_t_rep, bp_apd90, bp_apd50, bp_q0, bp_q1, bp_ca_dia, bp_ca_peak, bp_ca_amp;
  double bp_t1_cad90, bp_cad90, bp_t1_cad50, bp_cad50, bp_max_dvmdt, bp_vm_peak, bp_vm_dia, flags;

  enum : int { F_CA90_INC = 1, F_CAD90 = 2, F_CA50_INC = 4, F_CAD50 = 8 };

  CEP_HD void begin(double v0)
  {
    double *s = &v_prev;
    for (int k = 0; k < DR_SN; ++k) s[k] = 0.0;
    v_prev = v0;
    bp_max_dvmdt = v0;
    bp_vm_peak = v0;
  }

  // one output: k = outputs so far (1-based), t its time, acc = STATES[49..51] there,
  // dvdt_next = RATES[V] at t + dt with the states at t
  CEP_HD void sample(int k, double t, double v, double ca, double q, double inal, double ical, double dvdt_next,
                     double bcl, double dt, int n, double v90, double v50)
  {
    const int bs = (int)(bcl / dt);
    const int w_prev = bs * (n - 2) + (int)bcl, w_curr = bs * (n - 1) + (int)bcl, w_cadia = bs * (n - 2);
    const int imax = (int)((n * bcl) / dt);
    const double vp = v_prev;
    int fl = (int)flags;
    // drugs_herg
    if (k < imax) {
      if (vp > v90 && v <= v90) {
        dh_t_rep = t;
        dh_apd = dh_t_rep - dh_t_dep;
        dh_nrep += 1.0;
      }
      if (vp <= v90 && v > v90) dh_t_dep = t;
      if (k == w_prev) { dh_q0 = q; dh_n0 = inal; dh_c0 = ical; }
      else if (k == w_curr) { dh_q1 = q; dh_n1 = inal; dh_c1 = ical; }
    }
    // benchpatch
    const int i = k - 1;
    if (vp > v90 && v <= v90) {
      bp_t_rep = t;
      bp_apd90 = bp_t_rep - bp_t_dep;
    }
    if (vp <= v90 && v > v90) bp_t_dep = t;
    if (i == w_prev) bp_q0 = q;
    else if (i == w_curr) bp_q1 = q;
    if (i == w_cadia) bp_ca_dia = ca;
    if (t > bcl * (n - 2) && t < bcl * (n - 1) && bp_ca_peak < ca) {
      bp_ca_peak = ca;
      bp_ca_amp = bp_ca_peak - bp_ca_dia;
    }
    if (t > bcl * (n - 1)) {
      const double l90 = (0.1 * bp_ca_amp) + bp_ca_dia, l50 = (0.5 * bp_ca_amp) + bp_ca_dia;
      if (ca < l90 && (fl & F_CA90_INC) && !(fl & F_CAD90)) {
        bp_cad90 = t - bp_t1_cad90;
        fl |= F_CAD90;
      } else if (ca >= l90 && !(fl & F_CA90_INC)) {
        bp_t1_cad90 = t;
        fl |= F_CA90_INC;
      }
      if (ca < l50 && (fl & F_CA50_INC) && !(fl & F_CAD50)) {
        bp_cad50 = t - bp_t1_cad50;
        fl |= F_CAD50;
      } else if (ca >= l50 && !(fl & F_CA50_INC)) {
        bp_t1_cad50 = t;
        fl |= F_CA50_INC;
      }
      if (dvdt_next > bp_max_dvmdt) bp_max_dvmdt = dvdt_next;
      if (v > bp_vm_peak) bp_vm_peak = v;
      if (vp > v50 && v <= v50) {
        bp_t_rep = t;
        bp_apd50 = bp_t_rep - bp_t_dep;
      }
      if (vp <= v50 && v > v50) bp_t_dep = t;
    }
    flags = (double)fl;
    bp_vm_dia = v;
    v_prev = v;
  }

  CEP_HD void record(double *out) const
  {
    out[0] = dh_apd; out[1] = dh_t_dep; out[2] = dh_t_rep; out[3] = dh_nrep;
    out[4] = (dh_q1 - dh_q0) / 1000.0; out[5] = dh_n1 - dh_n0; out[6] = dh_c1 - dh_c0;
    out[7] = (bp_q1 - bp_q0) / 1000.0; out[8] = bp_max_dvmdt; out[9] = bp_vm_peak; out[10] = bp_vm_dia;
    out[11] = bp_apd90; out[12] = bp_apd50; out[13] = bp_ca_peak; out[14] = bp_ca_dia; out[15] = bp_cad90;
    out[16] = bp_cad50;
  }

  CEP_D void load(const double *d, int pitch, int c)
  {
    double *s = &v_prev;
    for (int k = 0; k < DR_SN; ++k) s[k] = d[(size_t)k * pitch + c];
  }
  CEP_D void store(double *d, int pitch, int c) const
  {
    const double *s = &v_prev;
    for (int k = 0; k < DR_SN; ++k) d[(size_t)k * pitch + c] = s[k];
  }
};
static_assert(sizeof(DriverRules) == DR_SN * sizeof(double), "state layout");

// population_tn2006's rules, one lane per cell (registers inside a launch, SoA between launches)
struct TnRules {
  double v_prev, t_dep, t_rep, apd, sum_ca, m;

  CEP_HD void sample(double tcurr, double v, double ca, double t_last, double v90)
  {
    if (v_prev > v90 && v <= v90) {
      t_rep = tcurr;
      apd = t_rep - t_dep;
    }
    if (v_prev <= v90 && v > v90) t_dep = tcurr;
    if (tcurr >= t_last) {
      sum_ca += ca;
      m += 1.0;
    }
    v_prev = v;
  }
  CEP_HD void record(double *out) const
  {
    out[0] = apd;
    out[1] = sum_ca / (m - 1.0);
  }
};

}  // namespace cep
