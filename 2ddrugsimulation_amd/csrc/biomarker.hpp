// biomarker.hpp -- per-cell on-device biomarker state machine.
//
// Restates the 14-biomarker extraction of CEPCell_NEO/modules/cipa_drug_bench.cpp:210-381
// as a per-sample update so it runs inside the integration kernels: rectangle-rule qNet /
// INaL / ICaL AUC (:225-227), per-beat Ca trace (:230-235) scanned at the beat boundary
// for CaD50/CaD90 (:238-263), steepest-repolarisation beat selection (:264-278), Vm peak
// sampled at iout % (int)((BCL*pace + stim_start + 6)/dt) == 0, repol30/50/90, dV/dt in
// [repol90, repol30], APD50/APD90 as the last time V sits in (level-2, level), Ca peak and
// the 50/90 % Ca levels (:294-334).  The window is the trailing `nwin` beats of the
// job (the reference hard-codes 250 and requires more than 250 paces).
// The Ca trace of the current beat lives in HBM (ctrace, SoA); everything else is a few
// dozen scalars per cell carried in registers inside a launch and in HBM between launches.
#pragma once
#include <cstddef>
#include <type_traits>
#include "common.hpp"

namespace cep {

struct BmResult {   // one beat's (or the selected beat's) biomarkers
  double qnet, inal_auc, ical_auc, dvmdt_repol, dvmdt_max, vm_peak, vm_valley, vm_dia, apd90,
      apd50, ca_peak, ca_valley, ca_dia, cad90, cad50;
};
constexpr int BM_RES_N = 15;
// persistent scalar state (besides the two results): doubles, then ints
constexpr int BM_D_N = 14;
constexpr int BM_I_N = 8;
constexpr int BM_TOTAL_D = BM_D_N + 2 * BM_RES_N;   // doubles stored per cell

// Optional time series of the selected (steepest) beat, for the reference's vmcheck / ca_i /
// ires / max_dvdt_repol files (CEPCell_NEO/modules/cipa_drug_bench.cpp:168-199,230-235,262,
// 338-350).  Samples (t, Vm, Ca_i, INaL, ICaL, Ito, IKr, IKs, IK1) at the Ca-trace points go
// to one of two slots per cell; when a beat becomes the selected one its slot is kept and
// the next beat writes the other, so no trace is copied.  Pointers are this cell's (offset
// by the cell index, strided by the pitch); t == nullptr: off.
struct TsView {
  double *t, *vm, *ca, *ires;   // [2][cap], [2][cap], [2][cap], [2][6][cap]
  int *rp;                      // [nwin] paces of the dvmdt_repol history
  double *rv;                   // [nwin] their dvmdt_repol
  int cap, nwin;
};

CEP_D void bm_res_init(BmResult &r, double v, double ca)
{
  r.qnet = 0.; r.inal_auc = 0.; r.ical_auc = 0.; r.dvmdt_repol = -999; r.dvmdt_max = -999;
  r.vm_peak = -999; r.vm_valley = v; r.vm_dia = -999; r.apd90 = 0.; r.apd50 = 0.;
  r.ca_peak = -999; r.ca_valley = ca; r.ca_dia = -999; r.cad90 = 0.; r.cad50 = 0.;
}

// The selection key of the selected beat (its dvmdt_repol), for machines that keep the
// selected beat's record in HBM.
struct BmKey { double dvmdt_repol; };

// BEST_IN_HBM = false: the selected beat's record is a member, like the rest of the state.
// BEST_IN_HBM = true: the record is written to its persistent rows (slots BM_D_N + BM_RES_N
// onwards of the SoA below) when a beat is selected and by begin(); the struct keeps only the
// selection key.  This is for machines held in LDS, where the 15 doubles per cell count.
// `best_row` is the cell's first row of that record, d + c + (BM_D_N + BM_RES_N) * pitch.
template <bool BEST_IN_HBM = false>
struct BiomarkerT {
  // configuration (per cell, from parameters)
  double bcl, dt, stim_start;
  int beat_steps, win_start, print_freq, cap;
  // persistent state
  double qnet, inal_auc, ical_auc, t_depol, repol30, repol50, repol90, t_ca_peak, ca_amp50,
      ca_amp90, cad50_prev, cad50_curr, cad90_prev, cad90_curr;
  int pace, steepest_pace, eligible, nca;
  int cur_slot, best_slot, best_n, nrep;   // time series (TsView): slot in use, selected slot
  BmResult temp;
  std::conditional_t<BEST_IN_HBM, BmKey, BmResult> best;

  // best = temp
  CEP_D void select_temp(double *best_row, int pitch)
  {
    if constexpr (BEST_IN_HBM) {
      best.dvmdt_repol = temp.dvmdt_repol;
      const double *tr = &temp.qnet;
      for (int k = 0; k < BM_RES_N; ++k) best_row[(size_t)k * pitch] = tr[k];
    } else {
      best = temp;
    }
  }

  CEP_D void begin(double v0, double ca0, double *best_row = nullptr, int pitch = 0)
  {
    bm_res_init(temp, v0, ca0);
    select_temp(best_row, pitch);
    qnet = inal_auc = ical_auc = 0.;
    pace = 0; steepest_pace = 0; eligible = 0; nca = 0;
    cur_slot = 0; best_slot = 0; best_n = 0; nrep = 0;
    t_depol = (bcl * pace) + stim_start;
    t_ca_peak = 0.; ca_amp50 = ca_amp90 = 0.;
    cad50_prev = cad50_curr = cad90_prev = cad90_curr = 0.;
    repol30 = repol50 = repol90 = 0.;
  }

  // ct/cv: this cell's Ca trace (time, value), strided by `pitch`
  // ires: INaL, ICaL, Ito, IKr, IKs, IK1 at this point (read only when ts.t is set)
  CEP_D void sample(int iout, double t, double v, double ca, double dvdt, double qsum, double inal,
                    double ical, double *ct, double *cv, int pitch, const TsView &ts = TsView{},
                    const double *ires = nullptr, double *best_row = nullptr)
  {
    qnet += qsum * dt;
    inal_auc += inal * dt;
    ical_auc += ical * dt;
    if (pace >= win_start && iout % print_freq == 0 && nca < cap) {
      ct[(size_t)nca * pitch] = t;
      cv[(size_t)nca * pitch] = ca;
      if (ts.t && nca < ts.cap) {
        const size_t o = ((size_t)cur_slot * ts.cap + nca) * pitch;
        ts.t[o] = t;
        ts.vm[o] = v;
        ts.ca[o] = ca;
        for (int j = 0; j < 6; ++j) ts.ires[((size_t)(cur_slot * 6 + j) * ts.cap + nca) * pitch] = ires[j];
      }
      nca++;
    }
    if (iout > 0 && iout % beat_steps == 0) {
      if (eligible && pace >= win_start) {
        for (int k = 0; k < nca; ++k) {
          const double tk = ct[(size_t)k * pitch], ck = cv[(size_t)k * pitch];
          if (tk < t_ca_peak) {
            if (ck < ca_amp50) cad50_prev = tk;
            if (ck < ca_amp90) cad90_prev = tk;
          } else {
            if (ck > ca_amp50) cad50_curr = tk;
            if (ck > ca_amp90) cad90_curr = tk;
          }
        }
        temp.cad50 = cad50_curr - cad50_prev;
        temp.cad90 = cad90_curr - cad90_prev;
        temp.qnet = qnet;
        temp.inal_auc = inal_auc;
        temp.ical_auc = ical_auc;
        temp.vm_dia = v;
        temp.ca_dia = ca;
        if (ts.rp && nrep < ts.nwin) {      // the reference's max_dvdt_repol line for this pace
          ts.rp[(size_t)nrep * pitch] = pace;
          ts.rv[(size_t)nrep * pitch] = temp.dvmdt_repol;
          nrep++;
        }
        if (temp.dvmdt_repol > best.dvmdt_repol) {
          steepest_pace = pace;
          select_temp(best_row, pitch);
          best_slot = cur_slot;
          best_n = nca;
        }
      }
      const int fin_slot = cur_slot, fin_n = nca;
      pace++;
      if (pace == win_start + 1) {
        select_temp(best_row, pitch);
        steepest_pace = pace - 1;
        best_slot = fin_slot;
        best_n = fin_n;
      }
      if (best_slot == fin_slot) cur_slot = 1 - fin_slot;   // keep the selected beat's samples
      qnet = inal_auc = ical_auc = 0.;
      if (pace >= win_start) {
        bm_res_init(temp, v, ca);
        nca = 0;
        t_ca_peak = t;
        t_depol = (bcl * pace) + stim_start;
        eligible = 0;
      }
    }
    if (pace >= win_start) {
      if (temp.dvmdt_max < dvdt) temp.dvmdt_max = dvdt;
      const int peak_idx = (int)(((bcl * pace) + (stim_start + 6.)) / dt);
      if (iout % peak_idx == 0) {
        temp.vm_peak = v;
        if (temp.vm_peak > 0.) {
          repol30 = temp.vm_peak - (0.3 * (temp.vm_peak - temp.vm_valley));
          repol50 = temp.vm_peak - (0.5 * (temp.vm_peak - temp.vm_valley));
          repol90 = temp.vm_peak - (0.9 * (temp.vm_peak - temp.vm_valley));
          eligible = 1;
        }
      } else if (eligible && t > (bcl * pace) + (stim_start + 6.)) {
        if (repol30 >= v && v >= repol90 && temp.dvmdt_repol < dvdt) temp.dvmdt_repol = dvdt;
        if (repol50 > v && v > repol50 - 2) temp.apd50 = t - t_depol;
        if (repol90 > v && v > repol90 - 2) temp.apd90 = t - t_depol;
        if (temp.ca_peak < ca) {
          temp.ca_peak = ca;
          ca_amp50 = temp.ca_peak - (0.5 * (temp.ca_peak - temp.ca_valley));
          ca_amp90 = temp.ca_peak - (0.9 * (temp.ca_peak - temp.ca_valley));
          t_ca_peak = t;
        }
      }
    }
  }

  // ---- persistence (SoA: slot k of cell c at d[k * pitch + c])
  CEP_D void load(const double *d, const int *iv, int pitch, int c)
  {
    const double *p = d + c;
    qnet = p[0]; inal_auc = p[(size_t)pitch]; ical_auc = p[2 * (size_t)pitch]; t_depol = p[3 * (size_t)pitch];
    repol30 = p[4 * (size_t)pitch]; repol50 = p[5 * (size_t)pitch]; repol90 = p[6 * (size_t)pitch];
    t_ca_peak = p[7 * (size_t)pitch]; ca_amp50 = p[8 * (size_t)pitch]; ca_amp90 = p[9 * (size_t)pitch];
    cad50_prev = p[10 * (size_t)pitch]; cad50_curr = p[11 * (size_t)pitch];
    cad90_prev = p[12 * (size_t)pitch]; cad90_curr = p[13 * (size_t)pitch];
    double *tr = &temp.qnet;
    for (int k = 0; k < BM_RES_N; ++k) tr[k] = p[(size_t)(BM_D_N + k) * pitch];
    if constexpr (BEST_IN_HBM) {
      best.dvmdt_repol = p[(size_t)(BM_D_N + BM_RES_N + offsetof(BmResult, dvmdt_repol) / 8) * pitch];
    } else {
      double *br = &best.qnet;
      for (int k = 0; k < BM_RES_N; ++k) br[k] = p[(size_t)(BM_D_N + BM_RES_N + k) * pitch];
    }
    const int *q = iv + c;
    pace = q[0]; steepest_pace = q[(size_t)pitch]; eligible = q[2 * (size_t)pitch]; nca = q[3 * (size_t)pitch];
    cur_slot = q[4 * (size_t)pitch]; best_slot = q[5 * (size_t)pitch]; best_n = q[6 * (size_t)pitch];
    nrep = q[7 * (size_t)pitch];
  }
  CEP_D void store(double *d, int *iv, int pitch, int c) const
  {
    double *p = d + c;
    p[0] = qnet; p[(size_t)pitch] = inal_auc; p[2 * (size_t)pitch] = ical_auc; p[3 * (size_t)pitch] = t_depol;
    p[4 * (size_t)pitch] = repol30; p[5 * (size_t)pitch] = repol50; p[6 * (size_t)pitch] = repol90;
    p[7 * (size_t)pitch] = t_ca_peak; p[8 * (size_t)pitch] = ca_amp50; p[9 * (size_t)pitch] = ca_amp90;
    p[10 * (size_t)pitch] = cad50_prev; p[11 * (size_t)pitch] = cad50_curr;
    p[12 * (size_t)pitch] = cad90_prev; p[13 * (size_t)pitch] = cad90_curr;
    const double *tr = &temp.qnet;
    for (int k = 0; k < BM_RES_N; ++k) p[(size_t)(BM_D_N + k) * pitch] = tr[k];
    if constexpr (!BEST_IN_HBM) {   // else the rows are written when a beat is selected
      const double *br = &best.qnet;
      for (int k = 0; k < BM_RES_N; ++k) p[(size_t)(BM_D_N + BM_RES_N + k) * pitch] = br[k];
    }
    int *q = iv + c;
    q[0] = pace; q[(size_t)pitch] = steepest_pace; q[2 * (size_t)pitch] = eligible; q[3 * (size_t)pitch] = nca;
    q[4 * (size_t)pitch] = cur_slot; q[5 * (size_t)pitch] = best_slot; q[6 * (size_t)pitch] = best_n;
    q[7 * (size_t)pitch] = nrep;
  }
};

using Biomarker = BiomarkerT<false>;

}  // namespace cep
