// tt06_model.hpp -- ten Tusscher 2006 (IK1Ko variant) endo/epi/M for gfx950.
//
// Equations: CEP4Cell/cellmodels/ten_tusscher_model_2006_IK1Ko_endo_units.cpp computeRates
// (:272-378) and solveAnalytical (:380-404); the epi and M files differ in initial states,
// g_to, g_Ks and the s-gate (epi and M share the epi s-gate).  STATES keep the reference's
// 19-entry layout (CEP7Cell/enums/enum_ten_tusscher_model_2006_IK1Ko.hpp).
// Precision: gating and currents in R; voltage/ion updates in double.
#pragma once
#include "common.hpp"

namespace cep {
namespace tt {

constexpr int NS = 19;
constexpr int NC = 53;
constexpr double Rg = 8314.472, Tk = 310, Fc = 96485.3415, Cm = 0.185, Vc = 0.016404;
constexpr double P_kna = 0.03, Ko = 5.4, Nao = 140, Cao = 2;
constexpr double K_mk = 1, K_mNa = 40, K_NaCa = 1000, K_sat = 0.1, alpha = 2.5, gmma = 0.35;
constexpr double Km_Ca = 1.38, Km_Nai = 87.5, K_pCa = 0.0005;
constexpr double k1_prime = 0.15, k2_prime = 0.045, k3 = 0.06, k4 = 0.005, EC = 1.5, max_sr = 2.5;
constexpr double min_sr = 1, V_rel = 0.102, V_xfer = 0.0038, K_up = 0.00025, V_leak = 0.00036;
constexpr double Vmax_up = 0.006375, Buf_c = 0.2, K_buf_c = 0.001, Buf_sr = 10, K_buf_sr = 0.3;
constexpr double Buf_ss = 0.4, K_buf_ss = 0.00025, V_sr = 0.001094, V_ss = 0.00005468;
constexpr double RTF = (Rg * Tk) / Fc;
constexpr double FRT = Fc / (Rg * Tk);
constexpr double sqrt_ko54 = 1.0;   // sqrt(Ko / 5.4)

enum : int {
  TP_CELLTYPE = 0, TP_GNA, TP_GK1, TP_GKR, TP_GKS, TP_GTO, TP_GCAL, TP_GBNA, TP_GBCA, TP_GPCA,
  TP_GPK, TP_PNAK, TP_STIM_START, TP_BCL, TP_STIM_DUR, TP_STIM_AMP, NP
};

struct Currents {
  double INa, ICaL, Ito, IKr, IKs, IK1, INaCa, INaK, IbNa, IbCa, IpK, IpCa, Istim;
};

template <typename R>
CEP_D R c(double x) { return (R)x; }

// Gate steady states (inf) and time constants for the 12 Rush-Larsen gates, in the order
// of the reference's solveAnalytical: f f2 fCass s r xr1 xr2 xs m h j d.
struct Gates {
  static constexpr int idx[12] = { 12, 13, 14, 15, 16, 4, 5, 6, 7, 8, 9, 11 };
};

template <typename R, typename PA>
CEP_D void gates(const double *__restrict__ y, const PA &P, R *inf, R *tau)
{
  using Mx = M<R>;
  const R V = (R)y[0];
  const R Cass = (R)y[10];
  const bool endo = P(TP_CELLTYPE) == 0.0;
  const R d27 = V + c<R>(27);
  inf[0] = Mx::rcp(c<R>(1) + Mx::exs((V + c<R>(20)), c<R>(1 / 7.0)));
  tau[0] = c<R>(1102.5) * Mx::exs(-d27 * d27, c<R>(1 / 225.0)) + Mx::dv(c<R>(200), c<R>(1) + Mx::exs((c<R>(13) - V), c<R>(0.1)))
         + Mx::dv(c<R>(180), c<R>(1) + Mx::exs((V + c<R>(30)), c<R>(0.1))) + c<R>(20);
  inf[1] = Mx::dv(c<R>(0.67), c<R>(1) + Mx::exs((V + c<R>(35)), c<R>(1 / 7.0))) + c<R>(0.33);
  tau[1] = c<R>(562) * Mx::exs(-d27 * d27, c<R>(1 / 240.0)) + Mx::dv(c<R>(31), c<R>(1) + Mx::exs((c<R>(25) - V), c<R>(0.1)))
         + Mx::dv(c<R>(80), c<R>(1) + Mx::exs((V + c<R>(30)), c<R>(0.1)));
  const R cq = Cass * c<R>(1 / 0.05), cq2 = c<R>(1) + cq * cq;
  inf[2] = Mx::dv(c<R>(0.6), cq2) + c<R>(0.4);
  tau[2] = Mx::dv(c<R>(80), cq2) + c<R>(2);
  if (endo) {
    inf[3] = Mx::rcp(c<R>(1) + Mx::exs((V + c<R>(28)), c<R>(0.2)));
    const R d67 = V + c<R>(67);
    tau[3] = c<R>(1000) * Mx::exs(-d67 * d67, c<R>(1 / 1000.0)) + c<R>(8);
  } else {
    inf[3] = Mx::rcp(c<R>(1) + Mx::exs((V + c<R>(20)), c<R>(0.2)));
    const R d45 = V + c<R>(45);
    tau[3] = c<R>(85) * Mx::exs(-d45 * d45, c<R>(1 / 320.0)) + Mx::dv(c<R>(5), c<R>(1) + Mx::exs((V - c<R>(20)), c<R>(0.2))) + c<R>(3);
  }
  inf[4] = Mx::rcp(c<R>(1) + Mx::exs((c<R>(20) - V), c<R>(1 / 6.0)));
  const R d40 = V + c<R>(40);
  tau[4] = c<R>(9.5) * Mx::exs(-d40 * d40, c<R>(1 / 1800.0)) + c<R>(0.8);
  inf[5] = Mx::rcp(c<R>(1) + Mx::exs((c<R>(-26) - V), c<R>(1 / 7.0)));
  tau[5] = Mx::dv(c<R>(450), c<R>(1) + Mx::exs((c<R>(-45) - V), c<R>(0.1))) * Mx::dv(c<R>(6), c<R>(1) + Mx::exs((V + c<R>(30)), c<R>(1 / 11.5)));
  inf[6] = Mx::rcp(c<R>(1) + Mx::exs((V + c<R>(88)), c<R>(1 / 24.0)));
  tau[6] = Mx::dv(c<R>(3), c<R>(1) + Mx::exs((c<R>(-60) - V), c<R>(0.05))) * Mx::dv(c<R>(1.12), c<R>(1) + Mx::exs((V - c<R>(60)), c<R>(0.05)));
  inf[7] = Mx::rcp(c<R>(1) + Mx::exs((c<R>(-5) - V), c<R>(1 / 14.0)));
  tau[7] = Mx::dv(c<R>(1400), Mx::sq(c<R>(1) + Mx::exs((c<R>(5) - V), c<R>(1 / 6.0))))
         * Mx::rcp(c<R>(1) + Mx::exs((V - c<R>(35)), c<R>(1 / 15.0))) + c<R>(80);
  const R em = c<R>(1) + Mx::exs((c<R>(-56.86) - V), c<R>(1 / 9.03));
  inf[8] = Mx::rcp(em * em);
  tau[8] = Mx::rcp(c<R>(1) + Mx::exs((c<R>(-60) - V), c<R>(0.2)))
         * (Mx::dv(c<R>(0.1), c<R>(1) + Mx::exs((V + c<R>(35)), c<R>(0.2))) + Mx::dv(c<R>(0.1), c<R>(1) + Mx::exs((V - c<R>(50)), c<R>(1 / 200.0))));
  const R eh = c<R>(1) + Mx::exs((V + c<R>(71.55)), c<R>(1 / 7.43));
  inf[9] = Mx::rcp(eh * eh);
  inf[10] = inf[9];
  R ah, bh, aj, bj;
  if (V < c<R>(-40)) {
    ah = c<R>(0.057) * Mx::exs(-(V + c<R>(80)), c<R>(1 / 6.8));
    bh = c<R>(2.7) * Mx::exs(V, c<R>(0.079)) + c<R>(310000.) * Mx::exs(V, c<R>(0.3485));
    aj = Mx::dv((c<R>(-25428.0) * Mx::exs(V, c<R>(0.2444)) - c<R>(6.948e-06) * Mx::exs(V, c<R>(-0.04391))) * (V + c<R>(37.78)),
                c<R>(1) + Mx::exs((V + c<R>(79.23)), c<R>(0.311)));
    bj = Mx::dv(c<R>(0.02424) * Mx::exs(V, c<R>(-0.01052)), c<R>(1) + Mx::exs((V + c<R>(40.14)), c<R>(-0.1378)));
  } else {
    ah = c<R>(0);
    bh = Mx::dv(c<R>(0.77), c<R>(0.13) * (c<R>(1) + Mx::exs((V + c<R>(10.66)), c<R>(-1 / 11.1))));
    aj = c<R>(0);
    bj = Mx::dv(c<R>(0.6) * Mx::exs(V, c<R>(0.057)), c<R>(1) + Mx::exs((V + c<R>(32)), c<R>(-0.1)));
  }
  tau[9] = Mx::rcp(ah + bh);
  tau[10] = Mx::rcp(aj + bj);
  inf[11] = Mx::rcp(c<R>(1) + Mx::exs((c<R>(-8) - V), c<R>(1 / 7.5)));
  tau[11] = (Mx::dv(c<R>(1.4), c<R>(1) + Mx::exs((c<R>(-35) - V), c<R>(1 / 13.0))) + c<R>(0.25))
          * Mx::dv(c<R>(1.4), c<R>(1) + Mx::exs((V + c<R>(5)), c<R>(0.2)))
          + Mx::rcp(c<R>(1) + Mx::exs((c<R>(50) - V), c<R>(0.05)));
}

// Currents and the non-gate rates.  f receives d/dt for V, K_i, Na_i, Ca_i, Ca_ss, Ca_SR,
// R_prime (indices 0,1,2,3,10,17,18); gate rates are written by rhs() below.
template <typename R, typename PA>
CEP_D void ionic(double t, const double *__restrict__ y, const PA &P, double *__restrict__ f, Currents *cur)
{
  using Mx = M<R>;
  const R V = (R)y[0];
  const R Ki = (R)y[1], Nai = (R)y[2], Cai = (R)y[3], Cass = (R)y[10], CaSR = (R)y[17], Rp = (R)y[18];
  const R INaK = Mx::dv(Mx::dv((R)P(TP_PNAK) * c<R>(Ko / (Ko + K_mk)) * Nai, Nai + c<R>(K_mNa)),
                        c<R>(1) + c<R>(0.1245) * Mx::exs(V, c<R>(-0.1 * FRT)) + c<R>(0.0353) * Mx::exs(V, c<R>(-FRT)));
  const R ENa = c<R>(RTF) * Mx::lg(Mx::dv(c<R>(Nao), Nai));
  const R m = (R)y[7];
  const R INa = (R)P(TP_GNA) * m * m * m * (R)y[8] * (R)y[9] * (V - ENa);
  const R IbNa = (R)P(TP_GBNA) * (V - ENa);
  const R egv = Mx::exs(V, c<R>(gmma * FRT)), eg1 = Mx::ex(c<R>((gmma - 1.0) * FRT) * V);
  const R INaCa = Mx::dv(c<R>(K_NaCa) * (egv * Nai * Nai * Nai * c<R>(Cao) - eg1 * c<R>(Nao * Nao * Nao) * Cai * c<R>(alpha)),
                         c<R>((Km_Nai * Km_Nai * Km_Nai + Nao * Nao * Nao) * (Km_Ca + Cao)) * (c<R>(1) + c<R>(K_sat) * eg1));
  const R EK = c<R>(RTF) * Mx::lg(Mx::dv(c<R>(Ko), Ki));
  const R dvk = V - EK;
  const R aK1 = Mx::dv(c<R>(0.1), c<R>(1) + Mx::exs((dvk - c<R>(200)), c<R>(0.06)));
  const R bK1 = Mx::dv(c<R>(3) * Mx::exs((dvk + c<R>(100)), c<R>(0.0002)) + Mx::exs((dvk - c<R>(10)), c<R>(0.1)),
                       c<R>(1) + Mx::exs(dvk, c<R>(-0.5)));
  const R IK1 = (R)P(TP_GK1) * Mx::dv(aK1, aK1 + bK1) * c<R>(sqrt_ko54) * dvk;
  const R Ito = (R)P(TP_GTO) * (R)y[16] * (R)y[15] * dvk;
  const R IKr = (R)P(TP_GKR) * c<R>(sqrt_ko54) * (R)y[4] * (R)y[5] * dvk;
  const R EKs = c<R>(RTF) * Mx::lg(Mx::dv(c<R>(Ko + P_kna * Nao), Ki + c<R>(P_kna) * Nai));
  const R xs = (R)y[6];
  const R IKs = (R)P(TP_GKS) * xs * xs * (V - EKs);
  // ICaL: GHK form.  FP64 keeps the reference's expression, 0/0 at exactly V = 15 mV
  // included (:327).  In FP32, e15 - 1 rounds to 0 for |V - 15| below ~1e-6 mV, far more
  // often than V hits 15 exactly in FP64, so the FP32 path evaluates (V-15)/(e15-1) by its
  // series 1/(2FRT) (1 - x/2), x = 2FRT (V-15), for |x| < 1e-4.
  const R v15 = V - c<R>(15);
  const R x15 = c<R>(2.0 * FRT) * v15;
  const R e15 = Mx::ex(x15);
  R ICaL;
  if constexpr (sizeof(R) == sizeof(float)) {
    const R q = (fabsf(x15) < 1e-4f) ? c<R>(1.0 / (2.0 * FRT)) * (c<R>(1) - c<R>(0.5) * x15)
                                     : Mx::dv(v15, e15 - c<R>(1));
    ICaL = (R)P(TP_GCAL) * (R)y[11] * (R)y[12] * (R)y[13] * (R)y[14] * c<R>(4.0 * Fc * FRT) * q
           * (c<R>(0.25) * Cass * e15 - c<R>(Cao));
  } else {
    ICaL = Mx::dv((R)P(TP_GCAL) * (R)y[11] * (R)y[12] * (R)y[13] * (R)y[14] * c<R>(4.0 * Fc * FRT) * v15
                      * (c<R>(0.25) * Cass * e15 - c<R>(Cao)),
                  e15 - c<R>(1));
  }
  const R ECa = c<R>(0.5 * RTF) * Mx::lg(Mx::dv(c<R>(Cao), Cai));
  const R IbCa = (R)P(TP_GBCA) * (V - ECa);
  const R IpK = Mx::dv((R)P(TP_GPK) * dvk, c<R>(1) + Mx::exs((c<R>(25) - V), c<R>(1 / 5.98)));
  const R IpCa = Mx::dv((R)P(TP_GPCA) * Cai, Cai + c<R>(K_pCa));
  const double bcl = P(TP_BCL);
  const double tm = t - floor(t / bcl) * bcl;
  const double st0 = P(TP_STIM_START);
  const double Istim = (tm >= st0 && tm <= st0 + P(TP_STIM_DUR)) ? -P(TP_STIM_AMP) : 0.0;
  // Ca dynamics (in double: fluxes of small concentrations)
  const double cai = y[3], cass = y[10], casr = y[17], rp = y[18];
  const double i_up = Vmax_up / (1.0 + (K_up * K_up) / (cai * cai));
  const double i_leak = V_leak * (casr - cai);
  const double i_xfer = V_xfer * (cass - cai);
  const double kcasr = max_sr - (max_sr - min_sr) / (1.0 + (EC / casr) * (EC / casr));
  const double k1 = k1_prime / kcasr, k2 = k2_prime * kcasr;
  const double Oo = (k1 * cass * cass * rp) / (k3 + k1 * cass * cass);
  const double i_rel = V_rel * Oo * (casr - cass);
  const double bufc = 1.0 / (1.0 + (Buf_c * K_buf_c) / ((cai + K_buf_c) * (cai + K_buf_c)));
  const double bufsr = 1.0 / (1.0 + (Buf_sr * K_buf_sr) / ((casr + K_buf_sr) * (casr + K_buf_sr)));
  const double bufss = 1.0 / (1.0 + (Buf_ss * K_buf_ss) / ((cass + K_buf_ss) * (cass + K_buf_ss)));
  const double dINa = INa, dIbNa = IbNa, dINaK = INaK, dINaCa = INaCa, dIK1 = IK1, dIto = Ito;
  const double dIKr = IKr, dIKs = IKs, dICaL = ICaL, dIbCa = IbCa, dIpK = IpK, dIpCa = IpCa;
  f[0] = -(dIK1 + dIto + dIKr + dIKs + dICaL + dINaK + dINa + dIbNa + dINaCa + dIbCa + dIpK + dIpCa + Istim);
  f[2] = (-(dINa + dIbNa + 3.0 * dINaK + 3.0 * dINaCa) / (Vc * Fc)) * Cm;
  f[1] = (-((dIK1 + dIto + dIKr + dIKs + dIpK + Istim) - 2.0 * dINaK) / (Vc * Fc)) * Cm;
  f[3] = bufc * ((((i_leak - i_up) * V_sr) / Vc + i_xfer) - ((dIbCa + dIpCa) - 2.0 * dINaCa) * Cm / (2.0 * Vc * Fc));
  f[18] = -k2 * cass * rp + k4 * (1.0 - rp);
  f[17] = bufsr * (i_up - (i_rel + i_leak));
  f[10] = bufss * ((-dICaL * Cm / (2.0 * V_ss * Fc) + (i_rel * V_sr) / V_ss) - (i_xfer * Vc) / V_ss);
  (void)Rp; (void)CaSR;
  if (cur) {
    cur->INa = dINa; cur->ICaL = dICaL; cur->Ito = dIto; cur->IKr = dIKr; cur->IKs = dIKs;
    cur->IK1 = dIK1; cur->INaCa = dINaCa; cur->INaK = dINaK; cur->IbNa = dIbNa; cur->IbCa = dIbCa;
    cur->IpK = dIpK; cur->IpCa = dIpCa; cur->Istim = Istim;
  }
}

// full RHS (RATES) for parity with computeRates
template <typename R, typename PA>
CEP_D void rhs(double t, const double *__restrict__ y, double *__restrict__ f, const PA &P, Currents *cur)
{
  R inf[12], tau[12];
  gates<R>(y, P, inf, tau);
  ionic<R>(t, y, P, f, cur);
#pragma unroll
  for (int g = 0; g < 12; ++g) {
    const int i = Gates::idx[g];
    f[i] = ((double)inf[g] - y[i]) / (double)tau[g];
  }
}

// One Rush-Larsen step (solveAnalytical): exact exponential update of the 12 gates with
// inf/tau evaluated at the step start, forward Euler for V and the ion concentrations.
template <typename R, typename PA>
CEP_D void rush_larsen(double t, double dt, double *__restrict__ y, const PA &P, double &dvdt, Currents *cur)
{
  using Mx = M<R>;
  R inf[12], tau[12];
  double f[19];
  gates<R>(y, P, inf, tau);
  ionic<R>(t, y, P, f, cur);
  dvdt = f[0];
#pragma unroll
  for (int g = 0; g < 12; ++g) {
    const int i = Gates::idx[g];
    if constexpr (sizeof(R) == sizeof(float)) {
      // FP32 rates: exp(-x) rounded to FP32 is off by up to 6e-8 near 1, which for a slow gate
      // (x = dt/tau ~ 1e-4 at dt 0.02 ms) is ~1e-3 of the step's change, with the same sign at
      // every step.  The change is formed from 1 - exp(-x) at its own relative accuracy instead:
      // the Taylor series to x^4 below x = 1/32 (truncation x^4/120 < 1e-8 of it).
      const R x = (R)dt * Mx::rcp(tau[g]);
      const R d = x < (R)0.03125 ? x * ((R)1 - x * ((R)0.5 - x * ((R)(1.0 / 6.0) - x * (R)(1.0 / 24.0))))
                                 : (R)1 - Mx::ex(-x);
      y[i] += ((double)inf[g] - y[i]) * (double)d;
    } else {
      const R e = Mx::ex(-(R)dt * Mx::rcp(tau[g]));
      y[i] = (y[i] - (double)inf[g]) * (double)e + (double)inf[g];
    }
  }
  y[0] += dt * f[0];
  y[18] += f[18] * dt;
  y[2] += f[2] * dt;
  y[1] += f[1] * dt;
  y[17] += f[17] * dt;
  y[10] += f[10] * dt;
  y[3] += f[3] * dt;
}

}  // namespace tt
}  // namespace cep
