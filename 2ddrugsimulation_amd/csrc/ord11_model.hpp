// ord11_model.hpp -- O'Hara-Rudy 2011 (Hodgkin-Huxley IKr, static Hill drug block) as driven by
// the DrugSimulation adaptive Rush-Larsen loop, for gfx950.
//
// Equations: DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp computeRates (:742-989), constants
// ___initConsts (:494-679), drug block ___applyDrugEffect (:469-485), Dutta scaling
// ___applyDutta (:460-467), Rush-Larsen step solveAnalytical (:991-1095) and the step-size
// rule set_time_step (:1097-1137).  STATES keep the reference's 41-entry layout
// (DrugSimulation/enums/enum_Ohara_Rudy_2011.hpp).
//
// The reference's initialisation stores the requested cell type and then resets it to 0, so
// every cell runs with the endocardial constants; those are compile-time here.  What varies per
// cell is the drug-blocked conductances, the four constants the Dutta scaling touches (it
// scales the slots whose indices equal the ALGEBRAIC indices of IKs, IKr, IK1, ICaL, INaL:
// Aff, GNaL, tjca, H; Gncx_b too, which no longer feeds anything), BCL and the stimulus.
#pragma once
#include "ord_model.hpp"

namespace cep {
namespace ord11 {

constexpr int NS = 41;    // STATES
constexpr int NC = 144;   // CONSTANTS (CellML layout, host side only)

// per-cell parameter slots
enum : int {
  P_GNA = 0, P_GNAL, P_GTO, P_PCA, P_GKR, P_GKS, P_GK1, P_AFF, P_TJCA, P_H, P_BCL, P_STIM_AMP, NP
};

// endocardial constants of ___initConsts (celltype == 0 branches)
constexpr double GNa = 75, GNaL_b = 0.0075, Gto_b = 0.02, PCa_b = 0.0001, GKr_b = 0.046;
constexpr double GKs_b = 0.0034, GK1_b = 0.1908, GKb = 0.003, Gncx = 0.0008, Pnak = 30;
constexpr double PCap = 1.1 * PCa_b, PCaNa = 0.00125 * PCa_b, PCaK = 0.0003574 * PCa_b;
constexpr double PCaNap = 0.00125 * PCap, PCaKp = 0.0003574 * PCap;
constexpr double cmdnmax = 0.05, upScale = 1.0;
constexpr double stim_start = 10.0, stim_dur = 0.5, stim_amp = -80.0;
// the Rush-Larsen update of fcafp in the reference divides the STATES index of d by tfcafp
constexpr double kFcafpNumerator = 25.0;

// currents the driver's features read (drug_sim.cpp:365-402)
struct Currents {
  double INa, INaL, Ito, ICaL, IKr, IKs, IK1, Istim;
};

// gate relaxations of one RHS evaluation for the states G0 .. NS-1 (the gates, nca and the two
// Jrel states): x' = inf + (x - inf) * exp(-dt / tau)
constexpr int G0 = 10, NG = NS - G0;
struct Gates {
  double inf[NG];
  double tau[NG];
  double nca_on;          // anca * k2n: the production term of dn/dt
};

template <typename R>
CEP_D R cc(double x) { return (R)x; }

// GHK driving factor A * U / (exp(U) - 1) as written in the model (no 0/0 guard) in double;
// float keeps a short series where exp(U) - 1 cancels
template <typename R>
CEP_D R ghk11(R A, R U, R eU);
template <>
CEP_D double ghk11<double>(double A, double U, double eU) { return (A * U) / (eU - 1.0); }
template <>
CEP_D float ghk11<float>(float A, float U, float eU)
{
  if (fabsf(U) < 1e-3f) return A * (1.0f - 0.5f * U + U * U * (1.0f / 12.0f));
  return A * __fdividef(U, eU - 1.0f);
}

// The voltage- and calcium-dependent gate kinetics of computeRates (:750-801): the steady
// states and time constants of gates 10-38 (nca: its production term and rate), in two
// independent groups: gate_rates_na, the INa / INaL / Ito gates 10-24 (V only), and
// gate_rates_ca, the ICaL / IKr / IKs / IK1 gates 25-38 (V, cass and jca).  Other wavefronts
// can compute them beside the currents.
template <typename R, typename PA, typename YV>
CEP_D void gate_rates_na(const YV &y, const PA &P, Gates &g)
{
  using Mx = M<R>;
  using namespace ord;
  (void)P;
  const R v = (R)y[0];
  const R hLss = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(87.61)) / cc<R>(7.488)));
  const R hLssp = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(93.81)) / cc<R>(7.488)));
  const R mss = Mx::rcp(cc<R>(1) + Mx::ex(-(v + cc<R>(mssV1)) / cc<R>(mssV2)));
  const R tm = Mx::rcp(cc<R>(mtD1) * Mx::ex((v + cc<R>(mtV1)) / cc<R>(mtV2)) + cc<R>(mtD2) * Mx::ex(-(v + cc<R>(mtV3)) / cc<R>(mtV4)));
  const R hss = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(hssV1)) / cc<R>(hssV2)));
  const R thf = Mx::rcp(cc<R>(1.432e-05) * Mx::ex(-(v + cc<R>(1.196)) / cc<R>(6.285)) + cc<R>(6.149) * Mx::ex((v + cc<R>(0.5096)) / cc<R>(20.27)));
  const R ths = Mx::rcp(cc<R>(0.009794) * Mx::ex(-(v + cc<R>(17.95)) / cc<R>(28.05)) + cc<R>(0.3343) * Mx::ex((v + cc<R>(5.73)) / cc<R>(56.66)));
  const R ass = Mx::rcp(cc<R>(1) + Mx::ex(-(v - cc<R>(14.34)) / cc<R>(14.82)));
  const R ta = Mx::dv(cc<R>(1.0515), Mx::rcp(cc<R>(1.2089) * (cc<R>(1) + Mx::ex(-(v - cc<R>(18.4099)) / cc<R>(29.3814))))
                                        + Mx::dv(cc<R>(3.5), cc<R>(1) + Mx::ex((v + cc<R>(100)) / cc<R>(29.3814))));
  const R jss = hss;
  const R tj = cc<R>(2.038) + Mx::rcp(cc<R>(0.02136) * Mx::ex(-(v + cc<R>(100.6)) / cc<R>(8.281)) + cc<R>(0.3052) * Mx::ex((v + cc<R>(0.9941)) / cc<R>(38.45)));
  const R assp = Mx::rcp(cc<R>(1) + Mx::ex(-(v - cc<R>(24.34)) / cc<R>(14.82)));
  const R hssp = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(89.1)) / cc<R>(6.086)));
  const R thsp = cc<R>(3) * ths;
  const R tjp = cc<R>(1.46) * tj;
  const R mLss = Mx::rcp(cc<R>(1) + Mx::ex(-(v + cc<R>(42.85)) / cc<R>(5.264)));
  const R tmL = tm;
  const R iss = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(43.94)) / cc<R>(5.711)));
  const R tiF = cc<R>(4.562) + Mx::rcp(cc<R>(0.3933) * Mx::ex(-(v + cc<R>(100)) / cc<R>(100)) + cc<R>(0.08004) * Mx::ex((v + cc<R>(50)) / cc<R>(16.59)));
  const R tiS = cc<R>(23.62) + Mx::rcp(cc<R>(0.001416) * Mx::ex(-(v + cc<R>(96.52)) / cc<R>(59.05)) + cc<R>(1.78e-08) * Mx::ex((v + cc<R>(114.1)) / cc<R>(8.079)));
  const R dti_develop = cc<R>(1.354) + Mx::dv(cc<R>(0.0001), Mx::ex((v - cc<R>(167.4)) / cc<R>(15.89)) + Mx::ex(-(v - cc<R>(12.23)) / cc<R>(0.2154)));
  const R dti_recover = cc<R>(1) - Mx::dv(cc<R>(0.5), cc<R>(1) + Mx::ex((v + cc<R>(70)) / cc<R>(20)));
  const R tiFp = dti_develop * dti_recover * tiF;
  const R tiSp = dti_develop * dti_recover * tiS;
  auto gate = [&](int i, R inf, R tau) {
    g.inf[i - G0] = (double)inf;
    g.tau[i - G0] = (double)tau;
  };
  gate(10, mss, tm); gate(11, hss, thf); gate(12, hss, ths); gate(13, jss, tj); gate(14, hssp, thsp);
  gate(15, jss, tjp); gate(16, mLss, tmL); gate(17, hLss, cc<R>(thL)); gate(18, hLssp, cc<R>(thLp));
  gate(19, ass, ta); gate(20, iss, tiF); gate(21, iss, tiS); gate(22, assp, ta); gate(23, iss, tiFp);
  gate(24, iss, tiSp);
}

template <typename R, typename PA, typename YV>
CEP_D void gate_rates_ca(const YV &y, const PA &P, Gates &g)
{
  using Mx = M<R>;
  using namespace ord;
  const R v = (R)y[0];
  const R cass = (R)y[2];
  const R dss = Mx::rcp(cc<R>(1) + Mx::ex(-(v + cc<R>(3.94)) / cc<R>(4.23)));
  const R td = cc<R>(0.6) + Mx::rcp(Mx::ex(cc<R>(-0.05) * (v + cc<R>(6))) + Mx::ex(cc<R>(0.09) * (v + cc<R>(14))));
  const R fss = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(19.58)) / cc<R>(3.696)));
  const R tff = cc<R>(7) + Mx::rcp(cc<R>(0.0045) * Mx::ex(-(v + cc<R>(20)) / cc<R>(10)) + cc<R>(0.0045) * Mx::ex((v + cc<R>(20)) / cc<R>(10)));
  const R tfs = cc<R>(1000) + Mx::rcp(cc<R>(3.5e-05) * Mx::ex(-(v + cc<R>(5)) / cc<R>(4)) + cc<R>(3.5e-05) * Mx::ex((v + cc<R>(5)) / cc<R>(6)));
  const R fcass = fss;
  const R jca = (R)y[30];
  const R km2n = jca * cc<R>(1);
  const R anca = Mx::rcp(Mx::dv(cc<R>(k2n), km2n) + [&] { const R q = cc<R>(1) + Mx::dv(cc<R>(Kmn), cass); const R q2 = q * q; return q2 * q2; }());
  const R xrss = Mx::rcp(cc<R>(1) + Mx::ex(-(v + cc<R>(8.337)) / cc<R>(6.789)));
  const R txrf = cc<R>(12.98) + Mx::rcp(cc<R>(0.3652) * Mx::ex((v - cc<R>(31.66)) / cc<R>(3.869)) + cc<R>(4.123e-05) * Mx::ex(-(v - cc<R>(47.78)) / cc<R>(20.38)));
  const R txrs = cc<R>(1.865) + Mx::rcp(cc<R>(0.06629) * Mx::ex((v - cc<R>(34.70)) / cc<R>(7.355)) + cc<R>(1.128e-05) * Mx::ex(-(v - cc<R>(29.74)) / cc<R>(25.94)));
  const R xs1ss = Mx::rcp(cc<R>(1) + Mx::ex(-(v + cc<R>(11.6)) / cc<R>(8.932)));
  const R txs1 = cc<R>(817.3) + Mx::rcp(cc<R>(0.0002326) * Mx::ex((v + cc<R>(48.28)) / cc<R>(17.8)) + cc<R>(0.001292) * Mx::ex(-(v + cc<R>(210)) / cc<R>(230)));
  const R xk1ss = Mx::rcp(cc<R>(1) + Mx::ex(-(v + cc<R>(2.5538 * ko + 144.59)) / cc<R>(1.5692 * ko + 3.8115)));
  const R txk1 = Mx::dv(cc<R>(122.2), Mx::ex(-(v + cc<R>(127.2)) / cc<R>(20.36)) + Mx::ex((v + cc<R>(236.8)) / cc<R>(69.33)));
  const R tfcaf = cc<R>(7) + Mx::rcp(cc<R>(0.04) * Mx::ex(-(v - cc<R>(4)) / cc<R>(7)) + cc<R>(0.04) * Mx::ex((v - cc<R>(4)) / cc<R>(7)));
  const R tfcas = cc<R>(100) + Mx::rcp(cc<R>(0.00012) * Mx::ex(-v / cc<R>(3)) + cc<R>(0.00012) * Mx::ex(v / cc<R>(7)));
  const R tffp = cc<R>(2.5) * tff;
  const R xs2ss = xs1ss;
  const R txs2 = Mx::rcp(cc<R>(0.01) * Mx::ex((v - cc<R>(50)) / cc<R>(20)) + cc<R>(0.0193) * Mx::ex(-(v + cc<R>(66.54)) / cc<R>(31)));
  const R tfcafp = cc<R>(2.5) * tfcaf;
  auto gate = [&](int i, R inf, R tau) {
    g.inf[i - G0] = (double)inf;
    g.tau[i - G0] = (double)tau;
  };
  gate(25, dss, td); gate(26, fss, tff); gate(27, fss, tfs); gate(28, fcass, tfcaf);
  gate(29, fcass, tfcas); gate(30, fcass, (R)P(P_TJCA)); gate(31, fss, tffp); gate(32, fcass, tfcafp);
  gate(34, xrss, txrf); gate(35, xrss, txrs); gate(36, xs1ss, txs1); gate(37, xs2ss, txs2);
  gate(38, xk1ss, txk1);
  // nca: dn/dt = anca k2n - n km2n; the step uses inf = anca k2n / km2n and rate km2n
  g.nca_on = (double)(anca * cc<R>(k2n));
  g.inf[33 - G0] = (double)(anca * cc<R>(k2n) / km2n);
  g.tau[33 - G0] = (double)km2n;     // a rate, not a time constant (see rl_step)
}

template <typename R, typename PA, typename YV>
CEP_D void gate_rates(const YV &y, const PA &P, Gates &g)
{
  gate_rates_na<R>(y, P, g);
  gate_rates_ca<R>(y, P, g);
}

// The currents of computeRates in two groups that share no intermediate result but CaMK's
// phosphorylated fraction and the GHK exponential (each group computes them): the Na and K
// currents (nak_currents) and the Ca currents, pumps, exchangers and fluxes (ca_currents).
// They hold the doubles the rates are formed from.
struct NaKCurrents {
  double INa, INaL, Ito, IKr, IKs, IK1, IKb, INab;
};
struct CaCurrents {
  double ICaL, ICaNa, ICaK, INaCa_i, INaCa_ss, INaK, IpCa, ICab;
  double JdiffK, JdiffNa, Jdiff, Jrel, Jup, Jtr, Bcass, Bcai, Bcajsr, f1;
  double Jrel_inf, tau_rel, Jrel_infp, tau_relp;
};

// INa, INaL, Ito, IKr, IKs, IK1, IKb, INab (:802-989)
template <typename R, typename PA, typename YV>
CEP_D void nak_currents(const YV &y, const PA &P, NaKCurrents &o)
{
  using Mx = M<R>;
  using namespace ord;   // shared physical constants (same values in both models)
  const R v = (R)y[0];
  const R camkt = (R)y[1], cass = (R)y[2], nai = (R)y[3], nass = (R)y[4], ki = (R)y[5];
  const R kss = (R)y[6], cansr = (R)y[7], cajsr = (R)y[8], cai = (R)y[9];
  (void)nass; (void)kss; (void)cansr; (void)cajsr;
  const R CaMKb = Mx::dv(cc<R>(CaMKo) * (cc<R>(1) - camkt), cc<R>(1) + Mx::dv(cc<R>(KmCaM), cass));
  const R vfrt = (v * cc<R>(Far)) / cc<R>(Rgas * Temp);
  const R e1 = Mx::ex(vfrt);
  const R CaMKa = CaMKb + camkt;
  const R fCaMKp = Mx::rcp(cc<R>(1) + Mx::dv(cc<R>(KmCaMK), CaMKa));

  // ---- K currents
  const R EK = cc<R>(RTF) * Mx::lg(Mx::dv(cc<R>(ko), ki));
  const R AiF = Mx::rcp(cc<R>(1) + Mx::ex((v - cc<R>(213.6)) / cc<R>(151.2)));
  const R AiS = cc<R>(1) - AiF;
  const R a = (R)y[19], iF = (R)y[20], iS = (R)y[21], ap = (R)y[22], iFp = (R)y[23], iSp = (R)y[24];
  const R ii = AiF * iF + AiS * iS;
  const R ip = AiF * iFp + AiS * iSp;
  const R Ito = (R)P(P_GTO) * (v - EK) * ((cc<R>(1) - fCaMKp) * a * ii + fCaMKp * ap * ip);
  const R Axrf = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(54.81)) / cc<R>(38.21)));
  const R Axrs = cc<R>(1) - Axrf;
  const R xr = Axrf * (R)y[34] + Axrs * (R)y[35];
  const R rkr = Mx::rcp(cc<R>(1) + Mx::ex((v + cc<R>(55)) / cc<R>(75))) * Mx::rcp(cc<R>(1) + Mx::ex((v - cc<R>(10)) / cc<R>(30)));
  const R IKr = (R)P(P_GKR) * cc<R>(1.0) /* sqrt(ko / 5.4) */ * xr * rkr * (v - EK);
  const R EKs = cc<R>(RTF) * Mx::lg(Mx::dv(cc<R>(ko + PKNa * nao), ki + cc<R>(PKNa) * nai));
  const R KsCa = cc<R>(1) + Mx::dv(cc<R>(0.6), cc<R>(1) + Mx::pw(Mx::dv(cc<R>(3.8e-05), cai), cc<R>(1.4)));
  const R xs1 = (R)y[36], xs2 = (R)y[37], xk1 = (R)y[38];
  const R IKs = (R)P(P_GKS) * KsCa * xs1 * xs2 * (v - EKs);
  const R rk1 = Mx::rcp(cc<R>(1) + Mx::ex(((v + cc<R>(105.8)) - cc<R>(2.6 * ko)) / cc<R>(9.493)));
  const R IK1 = (R)P(P_GK1) * cc<R>(2.3237900077244502) /* sqrt(ko) */ * rk1 * xk1 * (v - EK);

  const R IKb = cc<R>(GKb) * Mx::rcp(cc<R>(1) + Mx::ex(-(v - cc<R>(14.48)) / cc<R>(18.34))) * (v - EK);

  // ---- Na currents
  const R ENa = cc<R>(RTF) * Mx::lg(Mx::dv(cc<R>(nao), nai));
  const R m = (R)y[10], hf = (R)y[11], hs = (R)y[12], jj = (R)y[13], hsp = (R)y[14], jp = (R)y[15];
  const R mL = (R)y[16], hL = (R)y[17], hLp = (R)y[18];
  const R hh = cc<R>(Ahf) * hf + cc<R>(Ahs) * hs;
  const R hp = cc<R>(Ahf) * hf + cc<R>(Ahs) * hsp;
  const R INa = (R)P(P_GNA) * (v - ENa) * (m * m * m) * ((cc<R>(1) - fCaMKp) * hh * jj + fCaMKp * hp * jp);
  const R INaL = (R)P(P_GNAL) * (v - ENa) * mL * ((cc<R>(1) - fCaMKp) * hL + fCaMKp * hLp);

  const R INab = ghk11<R>(cc<R>(PNab * Far) * (nai * e1 - cc<R>(nao)), vfrt, e1);
  o.INa = (double)INa; o.INaL = (double)INaL; o.Ito = (double)Ito; o.IKr = (double)IKr;
  o.IKs = (double)IKs; o.IK1 = (double)IK1; o.IKb = (double)IKb; o.INab = (double)INab;
}

// ICaL, ICaNa, ICaK, INaK, INaCa (both compartments), IpCa, ICab, the fluxes, buffers, the
// Jrel relaxations and dCaMKt/dt (:802-989)
template <typename R, typename PA, typename YV>
CEP_D void ca_currents(const YV &y, const PA &P, CaCurrents &o)
{
  using Mx = M<R>;
  using namespace ord;   // shared physical constants (same values in both models)
  const R v = (R)y[0];
  const R camkt = (R)y[1], cass = (R)y[2], nai = (R)y[3], nass = (R)y[4], ki = (R)y[5];
  const R kss = (R)y[6], cansr = (R)y[7], cajsr = (R)y[8], cai = (R)y[9];
  const R jca = (R)y[30];
  const R CaMKb = Mx::dv(cc<R>(CaMKo) * (cc<R>(1) - camkt), cc<R>(1) + Mx::dv(cc<R>(KmCaM), cass));

  // ---- ICaL / ICaNa / ICaK
  const R Aff = (R)P(P_AFF);
  const R ff = (R)y[26], fs = (R)y[27], fcaf = (R)y[28], fcas = (R)y[29], ffp = (R)y[31], fcafp = (R)y[32];
  const R nca = (R)y[33], d = (R)y[25];
  // Afs = 1 - Aff was formed before any Dutta scaling of Aff
  const R f_ = Aff * ff + cc<R>(Afs) * fs;
  const R Afcaf = cc<R>(0.3) + Mx::dv(cc<R>(0.6), cc<R>(1) + Mx::ex((v - cc<R>(10)) / cc<R>(10)));
  const R Afcas = cc<R>(1) - Afcaf;
  const R fca = Afcaf * fcaf + Afcas * fcas;
  const R fp = Aff * ffp + cc<R>(Afs) * fs;
  const R fcap = Afcaf * fcafp + Afcas * fcas;
  const R vfrt = (v * cc<R>(Far)) / cc<R>(Rgas * Temp);
  const R e1 = Mx::ex(vfrt), e2 = Mx::ex(cc<R>(2) * vfrt);
  // the GHK factors written as A * U / (e^U - 1): U = 2 vfrt (Ca) or vfrt (Na, K)
  const R PhiCaL = ghk11<R>(cc<R>(2.0 * Far) * (cass * e2 - cc<R>(0.341 * cao)), cc<R>(2) * vfrt, e2);
  const R CaMKa = CaMKb + camkt;
  const R fCaMKp = Mx::rcp(cc<R>(1) + Mx::dv(cc<R>(KmCaMK), CaMKa));
  const R gnp = d * (f_ * (cc<R>(1) - nca) + jca * fca * nca);
  const R gp = d * (fp * (cc<R>(1) - nca) + jca * fcap * nca);
  const R ICaL = (cc<R>(1) - fCaMKp) * (R)P(P_PCA) * PhiCaL * gnp + fCaMKp * cc<R>(PCap) * PhiCaL * gp;
  const R q15 = Mx::dv(cc<R>(1.5), cajsr);
  const R q15_2 = q15 * q15, q15_4 = q15_2 * q15_2, q15_8 = q15_4 * q15_4;
  const R Jrel_inf = Mx::dv(cc<R>(a_rel) * -ICaL, cc<R>(1) + q15_8);
  const R tau_rel_t = Mx::dv(cc<R>(bt), cc<R>(1) + Mx::dv(cc<R>(0.0123), cajsr));
  const R tau_rel = tau_rel_t < cc<R>(0.001) ? cc<R>(0.001) : tau_rel_t;
  const R Jrel_infp = Mx::dv(cc<R>(a_relp) * -ICaL, cc<R>(1) + q15_8);
  const R tau_relp_t = Mx::dv(cc<R>(btp), cc<R>(1) + Mx::dv(cc<R>(0.0123), cajsr));
  const R tau_relp = tau_relp_t < cc<R>(0.001) ? cc<R>(0.001) : tau_relp_t;

  // ---- INaK (H scaled by the Dutta quirk enters P and b3)
  const R Hh = (R)P(P_H);
  const R Knao = cc<R>(Knao0) * Mx::ex((cc<R>(1.0 - delta) * v * cc<R>(Far)) / cc<R>(3.0 * Rgas * Temp));
  const R Knai = cc<R>(Knai0) * Mx::ex((cc<R>(delta) * v * cc<R>(Far)) / cc<R>(3.0 * Rgas * Temp));
  const R ro = cc<R>(1) + Mx::dv(cc<R>(nao), Knao), ro3 = ro * ro * ro;
  const R ri = cc<R>(1) + Mx::dv(nai, Knai), ri3 = ri * ri * ri;
  const R kki = cc<R>(1) + Mx::dv(ki, cc<R>(Kki)), kki2 = kki * kki;
  constexpr double kko = 1.0 + ko / Kko;
  const R den_o = (ro3 + cc<R>(kko * kko)) - cc<R>(1);
  const R den_i = (ri3 + kki2) - cc<R>(1);
  const R a3 = Mx::dv(cc<R>(k3p * (ko / Kko) * (ko / Kko)), den_o);
  const R Pp = Mx::dv(cc<R>(eP), cc<R>(1) + Mx::dv(Hh, cc<R>(Khp)) + Mx::dv(nai, cc<R>(Knap)) + Mx::dv(ki, cc<R>(Kxkur)));
  const R b3 = Mx::dv(cc<R>(k3m) * Pp * Hh, cc<R>(1.0 + MgATP / Kmgatp));
  const R nr = Mx::dv(nai, Knai);
  const R a1 = Mx::dv(cc<R>(k1p) * nr * nr * nr, den_i);
  const R orat = Mx::dv(cc<R>(nao), Knao);
  const R b2 = Mx::dv(cc<R>(k2m) * orat * orat * orat, den_o);
  const R kr = Mx::dv(ki, cc<R>(Kki));
  const R b4 = Mx::dv(cc<R>(k4m) * kr * kr, den_i);
  const R A2 = cc<R>(a2nak), A4 = cc<R>(a4nak), B1 = cc<R>(b1nak);
  const R x1 = A4 * a1 * A2 + b2 * b4 * b3 + A2 * b4 * b3 + b3 * a1 * A2;
  const R x2 = b2 * B1 * b4 + a1 * A2 * a3 + a3 * B1 * b4 + A2 * a3 * b4;
  const R x3 = A2 * a3 * A4 + b3 * b2 * B1 + b2 * B1 * A4 + a3 * A4 * B1;
  const R x4 = b4 * b3 * b2 + a3 * A4 * a1 + b2 * A4 * a1 + b3 * b2 * a1;
  const R xsum = x1 + x2 + x3 + x4;
  const R JnakNa = cc<R>(3) * (Mx::dv(x1, xsum) * a3 - Mx::dv(x2, xsum) * b3);
  const R JnakK = cc<R>(2) * (Mx::dv(x4, xsum) * B1 - Mx::dv(x3, xsum) * a1);
  const R INaK = cc<R>(Pnak) * (JnakNa + JnakK);
  const R JdiffK = (kss - ki) / cc<R>(2);
  const R PhiCaK = ghk11<R>(cc<R>(Far) * (cc<R>(0.75) * kss * e1 - cc<R>(0.75 * ko)), vfrt, e1);
  const R ICaK = (cc<R>(1) - fCaMKp) * cc<R>(PCaK) * PhiCaK * gnp + fCaMKp * cc<R>(PCaKp) * PhiCaK * gp;


  // ---- Na/Ca exchange
  const R hna = Mx::ex((cc<R>(qna) * v * cc<R>(Far)) / cc<R>(Rgas * Temp));
  const R hca = Mx::ex((cc<R>(qca) * v * cc<R>(Far)) / cc<R>(Rgas * Temp));
  const R h7 = cc<R>(1) + cc<R>(nao / kna3) * (cc<R>(1) + Mx::rcp(hna));
  const R h8 = Mx::dv(cc<R>(nao), cc<R>(kna3) * hna * h7);
  const R k3pp = h8 * cc<R>(wnaca);
  const R k8 = h8 * cc<R>(h11) * cc<R>(wna);
  const R k3 = Mx::rcp(h7) * cc<R>(wca) + k3pp;
  auto ncx = [&](R na, R ca, R frac) {
    const R h1 = cc<R>(1) + Mx::dv(na, cc<R>(kna3)) * (cc<R>(1) + hna);
    const R h2 = Mx::dv(na * hna, cc<R>(kna3) * h1);
    const R k4pp = h2 * cc<R>(wnaca);
    const R h4 = cc<R>(1) + Mx::dv(na, cc<R>(kna1)) * (cc<R>(1) + Mx::dv(na, cc<R>(kna2)));
    const R h5 = Mx::dv(na * na, h4 * cc<R>(kna1 * kna2));
    const R k7 = h5 * h2 * cc<R>(wna);
    const R k4 = Mx::dv(Mx::rcp(h1) * cc<R>(wca), hca) + k4pp;
    const R k6 = Mx::rcp(h4) * ca * cc<R>(kcaon);
    const R K1 = cc<R>(k1ncx), K2 = cc<R>(k2ncx), K5 = cc<R>(k5ncx);
    const R X1 = K2 * k4 * (k7 + k6) + K5 * k7 * (K2 + k3);
    const R X2 = K1 * k7 * (k4 + K5) + k4 * k6 * (K1 + k8);
    const R X3 = K1 * k3 * (k7 + k6) + k8 * k6 * (K2 + k3);
    const R X4 = K2 * k8 * (k4 + K5) + k3 * K5 * (K1 + k8);
    const R Xs = X1 + X2 + X3 + X4;
    const R E1 = Mx::dv(X1, Xs), E2 = Mx::dv(X2, Xs), E3 = Mx::dv(X3, Xs), E4 = Mx::dv(X4, Xs);
    const R JNa = (cc<R>(3) * (E4 * k7 - E1 * k8) + E3 * k4pp) - E2 * k3pp;
    const R JCa = E2 * K2 - E1 * K1;
    const R q = Mx::dv(cc<R>(KmCaAct), ca);
    const R allo = Mx::rcp(cc<R>(1) + q * q);
    return frac * cc<R>(Gncx) * allo * (JNa + cc<R>(2) * JCa);
  };
  const R INaCa_i = ncx(nai, cai, cc<R>(0.8));
  const R INaCa_ss = ncx(nass, cass, cc<R>(0.2));
  const R JdiffNa = (nass - nai) / cc<R>(2);
  const R PhiCaNa = ghk11<R>(cc<R>(Far) * (cc<R>(0.75) * nass * e1 - cc<R>(0.75 * nao)), vfrt, e1);
  const R ICaNa = (cc<R>(1) - fCaMKp) * cc<R>(PCaNa) * PhiCaNa * gnp + fCaMKp * cc<R>(PCaNap) * PhiCaNa * gp;
  const R IpCa = Mx::dv(cc<R>(GpCa) * cai, cc<R>(KmCap) + cai);
  const R ICab = ghk11<R>(cc<R>(PCab * 2.0 * Far) * (cai * e2 - cc<R>(0.341 * cao)), cc<R>(2) * vfrt, e2);
  const R Jdiff = (cass - cai) / cc<R>(0.2);
  const R Jrel = (cc<R>(1) - fCaMKp) * (R)y[39] + fCaMKp * (R)y[40];
  const R Bcass = Mx::rcp(cc<R>(1) + Mx::dv(cc<R>(BSRmax * KmBSR), (cc<R>(KmBSR) + cass) * (cc<R>(KmBSR) + cass))
                          + Mx::dv(cc<R>(BSLmax * KmBSL), (cc<R>(KmBSL) + cass) * (cc<R>(KmBSL) + cass)));
  const R Jupnp = Mx::dv(cc<R>(upScale * 0.004375) * cai, cai + cc<R>(0.00092));
  const R Jupp = Mx::dv(cc<R>(upScale * 2.75 * 0.004375) * cai, (cai + cc<R>(0.00092)) - cc<R>(0.00017));
  const R Jleak = cc<R>(0.0039375) * cansr / cc<R>(15);
  const R Jup = ((cc<R>(1) - fCaMKp) * Jupnp + fCaMKp * Jupp) - Jleak;
  const R Bcai = Mx::rcp(cc<R>(1) + Mx::dv(cc<R>(cmdnmax * kmcmdn), (cc<R>(kmcmdn) + cai) * (cc<R>(kmcmdn) + cai))
                         + Mx::dv(cc<R>(trpnmax * kmtrpn), (cc<R>(kmtrpn) + cai) * (cc<R>(kmtrpn) + cai)));
  const R Jtr = (cansr - cajsr) / cc<R>(100);
  const R Bcajsr = Mx::rcp(cc<R>(1) + Mx::dv(cc<R>(csqnmax * kmcsqn), (cc<R>(kmcsqn) + cajsr) * (cc<R>(kmcsqn) + cajsr)));

  o.ICaL = (double)ICaL; o.ICaNa = (double)ICaNa; o.ICaK = (double)ICaK; o.INaCa_i = (double)INaCa_i;
  o.INaCa_ss = (double)INaCa_ss; o.INaK = (double)INaK; o.IpCa = (double)IpCa; o.ICab = (double)ICab;
  o.JdiffK = (double)JdiffK; o.JdiffNa = (double)JdiffNa; o.Jdiff = (double)Jdiff; o.Jrel = (double)Jrel;
  o.Jup = (double)Jup; o.Jtr = (double)Jtr; o.Bcass = (double)Bcass; o.Bcai = (double)Bcai;
  o.Bcajsr = (double)Bcajsr;
  o.f1 = (double)(cc<R>(aCaMK) * CaMKb * (CaMKb + camkt) - cc<R>(bCaMK) * camkt);
  o.Jrel_inf = (double)Jrel_inf; o.tau_rel = (double)tau_rel;
  o.Jrel_infp = (double)Jrel_infp; o.tau_relp = (double)tau_relp;
}

// The rates f[0..9] of V, CaMKt and the concentrations, and the Jrel relaxations (39, 40),
// from the two current groups.  tissue_stim < 0: the single-cell periodic stimulus; 0 / 1:
// stimulus off / on (the TISSUE build).
template <typename R, typename PA, typename YV>
CEP_D void rates_from_currents(double t, const YV &y, double *f, const PA &P, Currents &cur, Gates &g,
                               const NaKCurrents &N, const CaCurrents &C, int tissue_stim = -1)
{
  using namespace ord;
  double Istim;
  if (tissue_stim >= 0) {
    Istim = tissue_stim ? P(P_STIM_AMP) : 0.0;
  } else {
    const double ts = t - stim_start, bcl = P(P_BCL);
    Istim = (t >= stim_start && ts - floor(ts / bcl) * bcl <= stim_dur) ? P(P_STIM_AMP) : 0.0;
  }
  cur.Istim = Istim;
  g.inf[39 - G0] = C.Jrel_inf;
  g.tau[39 - G0] = C.tau_rel;
  f[39] = (C.Jrel_inf - y[39]) / C.tau_rel;
  g.inf[40 - G0] = C.Jrel_infp;
  g.tau[40 - G0] = C.tau_relp;
  f[40] = (C.Jrel_infp - y[40]) / C.tau_relp;

  // ---- CaMK, concentrations, voltage (in double from the R-precision currents)
  f[1] = C.f1;
  const double dIstim = Istim;
  f[5] = (-((N.Ito + N.IKr + N.IKs + N.IK1 + N.IKb + dIstim) - 2.0 * C.INaK) * cm * Acap) / (Far * vmyo)
         + (C.JdiffK * vss) / vmyo;
  f[6] = (-C.ICaK * cm * Acap) / (Far * vss) - C.JdiffK;
  f[3] = (-(N.INa + N.INaL + 3.0 * C.INaCa_i + 3.0 * C.INaK + N.INab) * Acap * cm) / (Far * vmyo)
         + (C.JdiffNa * vss) / vmyo;
  f[4] = (-(C.ICaNa + 3.0 * C.INaCa_ss) * cm * Acap) / (Far * vss) - C.JdiffNa;
  f[0] = -(N.INa + N.INaL + N.Ito + C.ICaL + C.ICaNa + C.ICaK + N.IKr + N.IKs
           + N.IK1 + C.INaCa_i + C.INaCa_ss + C.INaK + N.INab + N.IKb + C.IpCa
           + C.ICab + dIstim);
  f[2] = C.Bcass * (((-(C.ICaL - 2.0 * C.INaCa_ss) * cm * Acap) / (2.0 * Far * vss)
                           + (C.Jrel * vjsr) / vss) - C.Jdiff);
  f[9] = C.Bcai * (((-((C.IpCa + C.ICab) - 2.0 * C.INaCa_i) * cm * Acap) / (2.0 * Far * vmyo)
                          - (C.Jup * vnsr) / vmyo) + (C.Jdiff * vss) / vmyo);
  f[7] = C.Jup - (C.Jtr * vjsr) / vnsr;
  f[8] = C.Bcajsr * (C.Jtr - C.Jrel);

  cur.INa = N.INa; cur.INaL = N.INaL; cur.Ito = N.Ito; cur.ICaL = C.ICaL; cur.IKr = N.IKr; cur.IKs = N.IKs;
  cur.IK1 = N.IK1;
}

// The rest of computeRates (:802-989): stimulus, currents, fluxes, the Jrel relaxations (39, 40:
// they depend on ICaL) and the rates f[0..9] of V, CaMKt and the concentrations.
template <typename R, typename PA, typename YV>
CEP_D void current_rates(double t, const YV &y, double *f, const PA &P, Currents &cur, Gates &g, int tissue_stim = -1)
{
  NaKCurrents N;
  CaCurrents C;
  nak_currents<R>(y, P, N);
  ca_currents<R>(y, P, C);
  rates_from_currents<R>(t, y, f, P, cur, g, N, C, tissue_stim);
}

// computeRates.  y, f: NS doubles.  g: steady states / time constants of the gates for the
// Rush-Larsen step.  tissue_stim < 0: the single-cell periodic stimulus; 0 / 1: stimulus
// off / on (the TISSUE build).
template <typename R, typename PA, typename YV>
CEP_D void rhs(double t, const YV &y, double *f, const PA &P, Currents &cur, Gates &g, int tissue_stim = -1)
{
  gate_rates<R>(y, P, g);
  current_rates<R>(t, y, f, P, cur, g, tissue_stim);
#pragma unroll
  for (int i = G0; i <= 38; ++i)
    if (i != 33) f[i] = (g.inf[i - G0] - y[i]) / g.tau[i - G0];
  f[33] = g.nca_on - y[33] * g.tau[33 - G0];
}

// solveAnalytical (:991-1095) in two parts, each element updated as the reference does:
// rl_gates, Rush-Larsen for the gates 10-38 (nca with its rate); rl_core, Rush-Larsen for the
// two Jrel states and forward Euler for V, CaMKt and the concentrations.  rl_step is both,
// with the rates and relaxations of the last rhs() call.
template <typename R, typename YV>
CEP_D void rl_gates_range(double dt, YV &y, const Gates &g, int lo, int hi)
{
  using Mx = M<R>;
#pragma unroll
  for (int i = G0; i <= 38; ++i) {
    if (i < lo || i > hi || i == 33) continue;
    const double inf = g.inf[i - G0];
    // fcafp: exp(-25 / tfcafp), the reference's use of the d gate's index in place of dt
    const double arg = (i == 32) ? -kFcafpNumerator / g.tau[i - G0] : -dt / g.tau[i - G0];
    y[i] = inf - (inf - y[i]) * (double)Mx::ex((R)arg);
  }
  if (lo <= 33 && 33 <= hi)
    y[33] = g.inf[33 - G0] - (g.inf[33 - G0] - y[33]) * (double)Mx::ex((R)(-g.tau[33 - G0] * dt));
}
// the gates of gate_rates_na (10-24) and of gate_rates_ca (25-38)
template <typename R, typename YV>
CEP_D void rl_gates_na(double dt, YV &y, const Gates &g) { rl_gates_range<R>(dt, y, g, G0, 24); }
template <typename R, typename YV>
CEP_D void rl_gates_ca(double dt, YV &y, const Gates &g) { rl_gates_range<R>(dt, y, g, 25, 38); }
template <typename R, typename YV>
CEP_D void rl_gates(double dt, YV &y, const Gates &g) { rl_gates_range<R>(dt, y, g, G0, 38); }

template <typename R>
CEP_D void rl_core(double dt, double *y, const double *f, const Gates &g)
{
  using Mx = M<R>;
#pragma unroll
  for (int i = 39; i <= 40; ++i) {
    const double inf = g.inf[i - G0];
    y[i] = inf - (inf - y[i]) * (double)Mx::ex((R)(-dt / g.tau[i - G0]));
  }
#pragma unroll
  for (int i = 0; i <= 9; ++i) y[i] += f[i] * dt;
}

template <typename R>
CEP_D void rl_step(double dt, double *y, const double *f, const Gates &g)
{
  rl_gates<R>(dt, y, g);
  rl_core<R>(dt, y, f, g);
}

// set_time_step: 0.005 ms inside the first time_point ms of every beat, else from |dV/dt|
CEP_HD double time_step(double t, double time_point, double max_dt, double bcl, double dvdt)
{
  double ts = 0.005;
  if (t <= time_point || (t - floor(t / bcl) * bcl) <= time_point) return ts;
  if (fabs(dvdt * ts) <= 0.2) {
    ts = fabs(0.8 / dvdt);
    if (ts < 0.005) ts = 0.005;
    else if (ts > max_dt) ts = max_dt;
  } else if (fabs(dvdt * ts) >= 0.8) {
    ts = fabs(0.2 / dvdt);
    while (fabs(dvdt * ts) >= 0.8 && 0.005 < ts && ts < max_dt) ts = ts / 10.0;
  }
  return ts;
}

}  // namespace ord11
}  // namespace cep
