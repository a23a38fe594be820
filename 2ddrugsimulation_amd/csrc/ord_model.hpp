// ord_model.hpp -- O'Hara-Rudy CiPA v1 2017 (Markov IKr + drug binding) for gfx950.
//
// Equations: CEP5Cell/cellmodels/ohara_rudy_cipa_v1_2017.cpp computeRates (:847-1111),
// constants/initial states ___initConsts (:545-804), drug block initConsts (:819-845).
// The state vector keeps the reference's 52-entry STATES layout (enum E_STATES_T of
// CEP7Cell/enums/enum_ohara_rudy_cipa_v1_2017.hpp) so rates/states cross the C ABI
// unchanged.  Constants shared by every cell are compile-time; the ~30 values that
// depend on cell type, drug block, population scaling and pacing are per-cell
// parameters (OP_* slots, SoA in HBM).
//
// Precision: ord_rhs<R> evaluates gating, Markov rates and currents in R (float or
// double); concentration/voltage rates are accumulated in double from the currents,
// and gate rates are formed as (double(inf) - state) so the cancellation near steady
// state happens in double.
#pragma once
#include "common.hpp"

namespace cep {
namespace ord {

constexpr int NS = 52;    // STATES
constexpr int NC = 206;   // CONSTANTS (CellML layout, host side only)

// ---- shared constants (___initConsts :547-748, derived :749-803)
constexpr double nao = 140, cao = 1.8, ko = 5.4, Rgas = 8314, Temp = 310, Far = 96485;
constexpr double Lcell = 0.01, rad = 0.0011;
constexpr double vcell = 1000.0 * 3.14 * rad * rad * Lcell;
constexpr double Ageo = 2.0 * 3.14 * rad * rad + 2.0 * 3.14 * rad * Lcell;
constexpr double Acap = 2.0 * Ageo;
constexpr double vmyo = 0.68 * vcell, vnsr = 0.0552 * vcell, vjsr = 0.0048 * vcell, vss = 0.02 * vcell;
constexpr double frt = Far / (Rgas * Temp);
constexpr double ffrt = Far * frt;
constexpr double RTF = (Rgas * Temp) / Far;
constexpr double KmCaMK = 0.15, aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015;
constexpr double kmcmdn = 0.00238, trpnmax = 0.07, kmtrpn = 0.0005, BSRmax = 0.047, KmBSR = 0.00087;
constexpr double BSLmax = 1.124, KmBSL = 0.0087, csqnmax = 10, kmcsqn = 0.8, cm = 1, PKNa = 0.01833;
constexpr double mssV1 = 39.57, mssV2 = 9.871, mtV1 = 11.64, mtV2 = 34.77, mtD1 = 6.765, mtD2 = 8.552;
constexpr double mtV3 = 77.42, mtV4 = 5.955, hssV1 = 82.9, hssV2 = 6.086, Ahf = 0.99, Ahs = 1.0 - Ahf;
constexpr double thL = 200, thLp = 3.0 * thL;
constexpr double Kmn = 0.002, k2n = 1000, tjca = 75.0, Aff = 0.6, Afs = 1.0 - Aff;
constexpr double bt = 4.75, a_rel = 0.5 * bt, btp = 1.25 * bt, a_relp = 0.5 * btp;
// INaCa
constexpr double kna1 = 15, kna2 = 5, kna3 = 88.12, kasymm = 12.5, wna = 6e4, wca = 6e4, wnaca = 5e3;
constexpr double kcaon = 1.5e6, kcaoff = 5e3, qna = 0.5224, qca = 0.167, KmCaAct = 150e-6;
constexpr double h10 = kasymm + 1.0 + (nao / kna1) * (1.0 + nao / kna2);
constexpr double h11 = (nao * nao) / (h10 * kna1 * kna2);
constexpr double h12 = 1.0 / h10;
constexpr double k1ncx = h12 * cao * kcaon, k2ncx = kcaoff, k5ncx = kcaoff;
// INaK
constexpr double k1p = 949.5, k1m = 182.4, k2p = 687.2, k2m = 39.4, k3p = 1899, k3m = 79300;
constexpr double k4p = 639, k4m = 40, Knai0 = 9.073, Knao0 = 27.78, delta = -0.155, Kki = 0.5;
constexpr double Kko = 0.3582, MgADP = 0.05, MgATP = 9.8, Kmgatp = 1.698e-7, H = 1e-7, eP = 4.2;
constexpr double Khp = 1.698e-7, Knap = 224, Kxkur = 292;
constexpr double b1nak = k1m * MgADP, a2nak = k2p;
constexpr double a4nak = ((k4p * MgATP) / Kmgatp) / (1.0 + MgATP / Kmgatp);
constexpr double PNab = 3.75e-10, PCab = 2.5e-8, GpCa = 0.0005, KmCap = 0.0005;
constexpr double txs1_max = 817.3;

// ---- per-cell parameter slots
enum : int {
  OP_CELLTYPE = 0, OP_GNA, OP_GNAL, OP_GTO, OP_PCA, OP_PCAP, OP_PCANA, OP_PCAK, OP_PCANAP,
  OP_PCAKP, OP_GKR, OP_GKS, OP_GK1, OP_GKB, OP_CMDNMAX, OP_UPSCALE, OP_GNCX, OP_PNAK,
  OP_KU, OP_BIND, OP_KT, OP_VHALF, OP_STIM_START, OP_STIM_END, OP_STIM_AMP, OP_BCL,
  OP_STIM_DUR, OP_CONC, OP_JREL_SCALE, OP_EPI, NP
};

// Markov IKr rate constants: A * exp(B * V) * q^((Temp-20)/10), the (A * Q10) products
// are formed on the host once (ohara_rudy_cipa_v1_2017.cpp:1060-1068).
struct MarkovTable {
  double AQ[14];
  double B[14];
};
// transition order: k1 k2 k3 k4 k11 k21 k31 k41 k51 k52 k53 k61 k62 k63
extern __constant__ MarkovTable g_markov;

// quantities the drivers read besides RATES (biomarkers, qNet, the CellML ALGEBRAIC slots)
struct Currents {
  double INa, INaL, Ito, ICaL, ICaNa, ICaK, IKr, IKs, IK1, INaCa_i, INaCa_ss, INaK, INab,
      IKb, IpCa, ICab, Istim, Jrel, Jup, CaMKa;
};

template <typename R>
CEP_D R c(double x) { return (R)x; }

// GHK-type driving factor A * U / (exp(U) - 1) with the reference's guard for |U| <= 1e-7
template <typename R>
CEP_D R ghk(R A, R U);
template <>
CEP_D double ghk<double>(double A, double U)
{
  return (-1.0e-07 <= U && U <= 1.0e-07) ? A * (1.0 - 0.5 * U) : (A * U) / (exp(U) - 1.0);
}
template <>
CEP_D float ghk<float>(float A, float U)
{
  // float: series near 0 where exp(U) - 1 cancels (|U| < 1e-3 -> error < U^4/720); elsewhere
  // the correctly rounded quotient, not the reciprocal-multiply of the other FP32 rates: this
  // driving force feeds ICaL, the largest plateau current, and the 1-ulp reciprocal here moved
  // a non-repolarising cell's Vm trace beyond the parity bar
  if (fabsf(U) < 1e-3f) return A * (1.0f - 0.5f * U + U * U * (1.0f / 12.0f));
  return A * __fdividef(U, __expf(U) - 1.0f);
}

// Stimulus of the single-cell drivers (SINGLE_CELL branch, :950-951)
CEP_D double stimulus(double t, double start, double end, double amp, double period, double dur)
{
  return (t >= start && t <= end && (t - start) - floor((t - start) / period) * period <= dur) ? amp : 0.0;
}

// sink for the gate relaxation rates of rhs() when the caller does not want them
struct NoGateRates {
  double d;
  CEP_D double &operator[](int) { return d; }
};

// Full RHS.  y/f: 52 doubles, indexable as y[i] / f[i] (register arrays or an LDS view).
// P(k): per-cell parameter k.  ga (optional): receives 1/tau of every Hodgkin-Huxley gate
// and of the two Jrel relaxations (f[i] = (inf - y[i]) * ga[i]), for Rush-Larsen steps.
template <typename R, typename PA, typename YV, typename FV, typename GA = NoGateRates>
CEP_D void rhs(double t, const YV &y, FV &&f, const PA &P, Currents *cur, GA &&ga = GA())
{
  using Mx = M<R>;
  const R v = (R)y[0];
  const double celltype = P(OP_CELLTYPE);
  const R camkt = (R)y[1], cass = (R)y[2], nai = (R)y[3], nass = (R)y[4], ki = (R)y[5];
  const R kss = (R)y[6], cajsr = (R)y[8], cai = (R)y[9];   // cansr (y[7]) enters in FP64 below

  // ---- HH gate steady states and time constants (:849-887)
  const R hLss = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(87.61)), c<R>(1 / 7.488)));
  const R hLssp = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(93.81)), c<R>(1 / 7.488)));
  const R mss = Mx::rcp(c<R>(1) + Mx::exs(-(v + c<R>(mssV1)), c<R>(1 / mssV2)));
  const R tm_inv = c<R>(mtD1) * Mx::exs((v + c<R>(mtV1)), c<R>(1 / mtV2)) + c<R>(mtD2) * Mx::exs(-(v + c<R>(mtV3)), c<R>(1 / mtV4));
  const R hss = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(hssV1)), c<R>(1 / hssV2)));
  const R thf_inv = c<R>(1.432e-05) * Mx::exs(-(v + c<R>(1.196)), c<R>(1 / 6.285)) + c<R>(6.149) * Mx::exs((v + c<R>(0.5096)), c<R>(1 / 20.27));
  const R ths_inv = c<R>(0.009794) * Mx::exs(-(v + c<R>(17.95)), c<R>(1 / 28.05)) + c<R>(0.3343) * Mx::exs((v + c<R>(5.73)), c<R>(1 / 56.66));
  const R ass = Mx::rcp(c<R>(1) + Mx::exs(-(v - c<R>(14.34)), c<R>(1 / 14.82)));
  const R ta = Mx::dv(c<R>(1.0515), Mx::rcp(c<R>(1.2089) * (c<R>(1) + Mx::exs(-(v - c<R>(18.4099)), c<R>(1 / 29.3814))))
                                        + Mx::dv(c<R>(3.5), c<R>(1) + Mx::exs((v + c<R>(100)), c<R>(1 / 29.3814))));
  const R dss = Mx::rcp(c<R>(1) + Mx::exs(-(v + c<R>(3.94)), c<R>(1 / 4.23)));
  const R td = c<R>(0.6) + Mx::rcp(Mx::exs((v + c<R>(6)), c<R>(-0.05)) + Mx::exs((v + c<R>(14)), c<R>(0.09)));
  const R fss = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(19.58)), c<R>(1 / 3.696)));
  const R tff = c<R>(7) + Mx::rcp(c<R>(0.0045) * Mx::exs(-(v + c<R>(20)), c<R>(0.1)) + c<R>(0.0045) * Mx::exs((v + c<R>(20)), c<R>(0.1)));
  const R tfs = c<R>(1000) + Mx::rcp(c<R>(3.5e-05) * Mx::exs(-(v + c<R>(5)), c<R>(0.25)) + c<R>(3.5e-05) * Mx::exs((v + c<R>(5)), c<R>(1 / 6.0)));
  const R jca = (R)y[30];
  const R anca = Mx::rcp(Mx::dv(c<R>(k2n), jca) + [&] { const R q = c<R>(1) + Mx::dv(c<R>(Kmn), cass); const R q2 = q * q; return q2 * q2; }());
  const R xs1ss = Mx::rcp(c<R>(1) + Mx::exs(-(v + c<R>(11.6)), c<R>(1 / 8.932)));
  const R txs1 = c<R>(txs1_max) + Mx::rcp(c<R>(0.0002326) * Mx::exs((v + c<R>(48.28)), c<R>(1 / 17.8)) + c<R>(0.001292) * Mx::exs(-(v + c<R>(210)), c<R>(1 / 230.0)));
  const R xk1ss = Mx::rcp(c<R>(1) + Mx::ex(-(v + c<R>(2.5538 * ko + 144.59)) * c<R>(1 / (1.5692 * ko + 3.8115))));
  const R txk1_inv = (Mx::exs(-(v + c<R>(127.2)), c<R>(1 / 20.36)) + Mx::exs((v + c<R>(236.8)), c<R>(1 / 69.33))) * c<R>(1 / 122.2);
  const R CaMKb = Mx::dv(c<R>(CaMKo) * (c<R>(1) - camkt), c<R>(1) + Mx::dv(c<R>(KmCaM), cass));
  const R tj = c<R>(2.038) + Mx::rcp(c<R>(0.02136) * Mx::exs(-(v + c<R>(100.6)), c<R>(1 / 8.281)) + c<R>(0.3052) * Mx::exs((v + c<R>(0.9941)), c<R>(1 / 38.45)));
  const R assp = Mx::rcp(c<R>(1) + Mx::exs(-(v - c<R>(24.34)), c<R>(1 / 14.82)));
  const R tfcaf = c<R>(7) + Mx::rcp(c<R>(0.04) * Mx::exs(-(v - c<R>(4)), c<R>(1 / 7.0)) + c<R>(0.04) * Mx::exs((v - c<R>(4)), c<R>(1 / 7.0)));
  const R tfcas = c<R>(100) + Mx::rcp(c<R>(0.00012) * Mx::exs(-v, c<R>(1 / 3.0)) + c<R>(0.00012) * Mx::exs(v, c<R>(1 / 7.0)));
  const R txs2_inv = c<R>(0.01) * Mx::exs((v - c<R>(50)), c<R>(0.05)) + c<R>(0.0193) * Mx::exs(-(v + c<R>(66.54)), c<R>(1 / 31.0));
  const R hssp = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(89.1)), c<R>(1 / 6.086)));
  const R mLss = Mx::rcp(c<R>(1) + Mx::exs(-(v + c<R>(42.85)), c<R>(1 / 5.264)));
  const R iss = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(43.94)), c<R>(1 / 5.711)));
  const R delta_epi = (celltype == 1.0) ? c<R>(1) - Mx::dv(c<R>(0.95), c<R>(1) + Mx::exs((v + c<R>(70)), c<R>(0.2))) : c<R>(1);
  const R tiF = (c<R>(4.562) + Mx::rcp(c<R>(0.3933) * Mx::exs(-(v + c<R>(100)), c<R>(0.01)) + c<R>(0.08004) * Mx::exs((v + c<R>(50)), c<R>(1 / 16.59)))) * delta_epi;
  const R tiS = (c<R>(23.62) + Mx::rcp(c<R>(0.001416) * Mx::exs(-(v + c<R>(96.52)), c<R>(1 / 59.05)) + c<R>(1.78e-08) * Mx::exs((v + c<R>(114.1)), c<R>(1 / 8.079)))) * delta_epi;
  const R dti_develop = c<R>(1.354) + Mx::dv(c<R>(0.0001), Mx::exs((v - c<R>(167.4)), c<R>(1 / 15.89)) + Mx::exs(-(v - c<R>(12.23)), c<R>(1 / 0.2154)));
  const R dti_recover = c<R>(1) - Mx::dv(c<R>(0.5), c<R>(1) + Mx::exs((v + c<R>(70)), c<R>(0.05)));
  const R dti = dti_develop * dti_recover;

  // ---- ICaL, ICaNa, ICaK (:888-896, :962-965, :1003-1006)
  const R ff = (R)y[26], fs = (R)y[27], fcaf = (R)y[28], fcas = (R)y[29], ffp = (R)y[31], fcafp = (R)y[32];
  const R nca = (R)y[33], d = (R)y[25];
  const R f_ = c<R>(Aff) * ff + c<R>(Afs) * fs;
  const R Afcaf = c<R>(0.3) + Mx::dv(c<R>(0.6), c<R>(1) + Mx::exs((v - c<R>(10)), c<R>(0.1)));
  const R Afcas = c<R>(1) - Afcaf;
  const R fca = Afcaf * fcaf + Afcas * fcas;
  const R fp = c<R>(Aff) * ffp + c<R>(Afs) * fs;
  const R fcap = Afcaf * fcafp + Afcas * fcas;
  const R vfrt = v * c<R>(frt);
  const R e1 = Mx::ex(vfrt), e2 = e1 * e1;
  const R U1 = c<R>(2.0 * frt) * v;
  const R PhiCaL = ghk<R>(Mx::dv(c<R>(4.0 * ffrt) * (cass * e2 - c<R>(0.341 * cao)), c<R>(2.0 * frt)), U1);
  const R CaMKa = CaMKb + camkt;
  const R fCaMKp = Mx::rcp(c<R>(1) + Mx::dv(c<R>(KmCaMK), CaMKa));   // same phosphorylation fraction used by ICaL/Ito/INa/INaL/Jrel/Jup
  const R gate_np = d * (f_ * (c<R>(1) - nca) + jca * fca * nca);
  const R gate_p = d * (fp * (c<R>(1) - nca) + jca * fcap * nca);
  const R ICaL = (c<R>(1) - fCaMKp) * (R)P(OP_PCA) * PhiCaL * gate_np + fCaMKp * (R)P(OP_PCAP) * PhiCaL * gate_p;
  const R U2 = c<R>(frt) * v;
  const R PhiCaNa = ghk<R>(Mx::dv(c<R>(0.75 * ffrt) * (nass * e1 - c<R>(nao)), c<R>(frt)), U2);
  const R PhiCaK = ghk<R>(Mx::dv(c<R>(0.75 * ffrt) * (kss * e1 - c<R>(ko)), c<R>(frt)), U2);
  const R ICaNa = (c<R>(1) - fCaMKp) * (R)P(OP_PCANA) * PhiCaNa * gate_np + fCaMKp * (R)P(OP_PCANAP) * PhiCaNa * gate_p;
  const R ICaK = (c<R>(1) - fCaMKp) * (R)P(OP_PCAK) * PhiCaK * gate_np + fCaMKp * (R)P(OP_PCAKP) * PhiCaK * gate_p;

  // ---- Jrel targets (:897-904)
  const R q15 = Mx::dv(c<R>(1.5), cajsr);
  const R q15_2 = q15 * q15, q15_4 = q15_2 * q15_2, q15_8 = q15_4 * q15_4;
  const R jscale = (R)P(OP_JREL_SCALE);
  const R Jrel_inf = jscale * Mx::dv(c<R>(a_rel) * -ICaL, c<R>(1) + q15_8);
  const R tau_rel_b = Mx::dv(c<R>(bt), c<R>(1) + Mx::dv(c<R>(0.0123), cajsr));
  const R tau_rel = tau_rel_b < c<R>(0.001) ? c<R>(0.001) : tau_rel_b;
  const R Jrel_infp = jscale * Mx::dv(c<R>(a_relp) * -ICaL, c<R>(1) + q15_8);
  const R tau_relp_b = Mx::dv(c<R>(btp), c<R>(1) + Mx::dv(c<R>(0.0123), cajsr));
  const R tau_relp = tau_relp_b < c<R>(0.001) ? c<R>(0.001) : tau_relp_b;

  // ---- K currents (:905-918)
  const R EK = c<R>(RTF) * Mx::lg(Mx::dv(c<R>(ko), ki));
  const R AiF = Mx::rcp(c<R>(1) + Mx::exs((v - c<R>(213.6)), c<R>(1 / 151.2)));
  const R AiS = c<R>(1) - AiF;
  const R a = (R)y[19], iF = (R)y[20], iS = (R)y[21], ap = (R)y[22], iFp = (R)y[23], iSp = (R)y[24];
  const R ii = AiF * iF + AiS * iS;
  const R ip = AiF * iFp + AiS * iSp;
  const R Ito = (R)P(OP_GTO) * (v - EK) * ((c<R>(1) - fCaMKp) * a * ii + fCaMKp * ap * ip);
  const R O = (R)y[38];
  const R IKr = (R)P(OP_GKR) * c<R>(1.0) /* sqrt(ko/5.4) */ * O * (v - EK);
  const R EKs = c<R>(RTF) * Mx::lg(Mx::dv(c<R>(ko + PKNa * nao), ki + c<R>(PKNa) * nai));
  const R KsCa = c<R>(1) + Mx::dv(c<R>(0.6), c<R>(1) + Mx::pw(Mx::dv(c<R>(3.8e-05), cai), c<R>(1.4)));
  const R xs1 = (R)y[44], xs2 = (R)y[45], xk1 = (R)y[46];
  const R IKs = (R)P(OP_GKS) * KsCa * xs1 * xs2 * (v - EKs);
  const R rk1 = Mx::rcp(c<R>(1) + Mx::exs((v + c<R>(105.8 - 2.6 * ko)), c<R>(1 / 9.493)));
  const R IK1 = (R)P(OP_GK1) * c<R>(2.3237900077244502) /* sqrt(ko) */ * rk1 * xk1 * (v - EK);

  // ---- INaK (:919-941)
  const R Knao = c<R>(Knao0) * Mx::ex(c<R>((1.0 - delta) / 3.0 * frt) * v);
  const R Knai = c<R>(Knai0) * Mx::exs(v, c<R>(delta / 3.0 * frt));
  const R ro = c<R>(1) + Mx::dv(c<R>(nao), Knao), ro3 = ro * ro * ro;
  const R ri = c<R>(1) + Mx::dv(nai, Knai), ri3 = ri * ri * ri;
  const R kki = c<R>(1) + ki * c<R>(1 / Kki), kki2 = kki * kki;
  constexpr double kko = 1.0 + ko / Kko;
  const R den_o = ro3 + c<R>(kko * kko) - c<R>(1);
  const R den_i = ri3 + kki2 - c<R>(1);
  const R a3 = Mx::dv(c<R>(k3p * (ko / Kko) * (ko / Kko)), den_o);
  const R Pnak_P = Mx::dv(c<R>(eP), c<R>(1 + H / Khp) + nai * c<R>(1 / Knap) + ki * c<R>(1 / Kxkur));
  const R b3 = c<R>(k3m * H / (1.0 + MgATP / Kmgatp)) * Pnak_P;
  const R nr = Mx::dv(nai, Knai);
  const R a1 = Mx::dv(c<R>(k1p) * nr * nr * nr, den_i);
  const R orat = Mx::dv(c<R>(nao), Knao);
  const R b2 = Mx::dv(c<R>(k2m) * orat * orat * orat, den_o);
  const R kr = ki * c<R>(1 / Kki);
  const R b4 = Mx::dv(c<R>(k4m) * kr * kr, den_i);
  const R A2 = c<R>(a2nak), A4 = c<R>(a4nak), B1 = c<R>(b1nak);
  const R x1 = A4 * a1 * A2 + b2 * b4 * b3 + A2 * b4 * b3 + b3 * a1 * A2;
  const R x2 = b2 * B1 * b4 + a1 * A2 * a3 + a3 * B1 * b4 + A2 * a3 * b4;
  const R x3 = A2 * a3 * A4 + b3 * b2 * B1 + b2 * B1 * A4 + a3 * A4 * B1;
  const R x4 = b4 * b3 * b2 + a3 * A4 * a1 + b2 * A4 * a1 + b3 * b2 * a1;
  const R xs_inv = Mx::rcp(x1 + x2 + x3 + x4);
  const R JnakNa = c<R>(3) * (x1 * xs_inv * a3 - x2 * xs_inv * b3);
  const R JnakK = c<R>(2) * (x4 * xs_inv * B1 - x3 * xs_inv * a1);
  const R INaK = (R)P(OP_PNAK) * (JnakNa + JnakK);

  const R IKb = (R)P(OP_GKB) * Mx::rcp(c<R>(1) + Mx::exs(-(v - c<R>(14.48)), c<R>(1 / 18.34))) * (v - EK);
  const double Istim = stimulus(t, P(OP_STIM_START), P(OP_STIM_END), P(OP_STIM_AMP), P(OP_BCL), P(OP_STIM_DUR));

  // ---- Na currents (:966-973)
  const R ENa = c<R>(RTF) * Mx::lg(Mx::dv(c<R>(nao), nai));
  const R m = (R)y[10], hf = (R)y[11], hs = (R)y[12], jj = (R)y[13], hsp = (R)y[14], jp = (R)y[15];
  const R mL = (R)y[16], hL = (R)y[17], hLp = (R)y[18];
  const R hh = c<R>(Ahf) * hf + c<R>(Ahs) * hs;
  const R hp = c<R>(Ahf) * hf + c<R>(Ahs) * hsp;
  const R INa = (R)P(OP_GNA) * (v - ENa) * (m * m * m) * ((c<R>(1) - fCaMKp) * hh * jj + fCaMKp * hp * jp);
  const R INaL = (R)P(OP_GNAL) * (v - ENa) * mL * ((c<R>(1) - fCaMKp) * hL + fCaMKp * hLp);

  // ---- INaCa_i and INaCa_ss (:974-1034): shared Na/Ca-independent terms
  const R hna = Mx::exs(v, c<R>(qna * frt));
  const R hca = Mx::exs(v, c<R>(qca * frt));
  const R h7 = c<R>(1) + c<R>(nao / kna3) * (c<R>(1) + Mx::rcp(hna));
  const R h8 = Mx::dv(c<R>(nao / kna3), hna * h7);
  const R k3pp = h8 * c<R>(wnaca);
  const R k8 = h8 * c<R>(h11 * wna);
  const R k3 = Mx::rcp(h7) * c<R>(wca) + k3pp;
  const R hca_inv = Mx::rcp(hca);
  auto ncx = [&](R na, R ca, R &JNa, R &JCa) {
    const R h1 = c<R>(1) + na * c<R>(1 / kna3) * (c<R>(1) + hna);
    const R h1i = Mx::rcp(h1);
    const R h2 = na * hna * c<R>(1 / kna3) * h1i;
    const R k4pp = h2 * c<R>(wnaca);
    const R h4 = c<R>(1) + na * c<R>(1 / kna1) * (c<R>(1) + na * c<R>(1 / kna2));
    const R h4i = Mx::rcp(h4);
    const R h5 = na * na * h4i * c<R>(1 / (kna1 * kna2));
    const R k7 = h5 * h2 * c<R>(wna);
    const R k4 = h1i * c<R>(wca) * hca_inv + k4pp;
    const R k6 = h4i * ca * c<R>(kcaon);
    const R K1 = c<R>(k1ncx), K2 = c<R>(k2ncx), K5 = c<R>(k5ncx);
    const R X1 = K2 * k4 * (k7 + k6) + K5 * k7 * (K2 + k3);
    const R X2 = K1 * k7 * (k4 + K5) + k4 * k6 * (K1 + k8);
    const R X3 = K1 * k3 * (k7 + k6) + k8 * k6 * (K2 + k3);
    const R X4 = K2 * k8 * (k4 + K5) + k3 * K5 * (K1 + k8);
    const R Xi = Mx::rcp(X1 + X2 + X3 + X4);
    const R E1 = X1 * Xi, E2 = X2 * Xi, E3 = X3 * Xi, E4 = X4 * Xi;
    JNa = (c<R>(3) * (E4 * k7 - E1 * k8) + E3 * k4pp) - E2 * k3pp;
    JCa = E2 * K2 - E1 * K1;
  };
  R JncxNa_i, JncxCa_i, JncxNa_s, JncxCa_s;
  ncx(nai, cai, JncxNa_i, JncxCa_i);
  ncx(nass, cass, JncxNa_s, JncxCa_s);
  const R alloi = Mx::rcp(c<R>(1) + [&] { const R q = Mx::dv(c<R>(KmCaAct), cai); return q * q; }());
  const R allos = Mx::rcp(c<R>(1) + [&] { const R q = Mx::dv(c<R>(KmCaAct), cass); return q * q; }());
  const R Gncx = (R)P(OP_GNCX);
  const R INaCa_i = c<R>(0.8) * Gncx * alloi * (JncxNa_i + c<R>(2) * JncxCa_i);
  const R INaCa_ss = c<R>(0.2) * Gncx * allos * (JncxNa_s + c<R>(2) * JncxCa_s);

  // ---- background / pump (:1000-1036)
  const R INab = ghk<R>(Mx::dv(c<R>(PNab * ffrt) * (nai * e1 - c<R>(nao)), c<R>(frt)), U2);
  const R IpCa = Mx::dv(c<R>(GpCa) * cai, c<R>(KmCap) + cai);
  const R ICab = ghk<R>(Mx::dv(c<R>(PCab * 4.0 * ffrt) * (cai * e2 - c<R>(0.341 * cao)), c<R>(2.0 * frt)), U1);

  // ---- Markov IKr with drug binding (:1059-1068)
  R k[14];
#pragma unroll
  for (int i = 0; i < 14; ++i) k[i] = (R)g_markov.AQ[i] * Mx::ex((R)g_markov.B[i] * v);
  const R k1 = k[0], k2 = k[1], k3m_ = k[2], k4m_ = k[3], k11 = k[4], k21 = k[5], k31 = k[6], k41 = k[7];
  const R k51 = k[8], k52 = k[9], k53 = k[10], k61 = k[11], k62 = k[12], k63 = k[13];
  const R IC1 = (R)y[34], IC2 = (R)y[35], C1 = (R)y[36], C2 = (R)y[37], IO = (R)y[39];
  const R IObound = (R)y[40], Obound = (R)y[41], Cbound = (R)y[42];
  const R Ku = (R)P(OP_KU), Kt = (R)P(OP_KT), bind = (R)P(OP_BIND);
  const R trap = Mx::dv(Kt, c<R>(1) + Mx::exs(-(v - (R)P(OP_VHALF)), c<R>(1 / 6.789)));
  const R unbindIO = Mx::dv(Ku * k53, k63);
  const R t_IC1_IC2 = k11 * IC1 - k21 * IC2, t_C1_IC1 = k51 * C1 - k61 * IC1;
  const R t_IC2_IO = k3m_ * IC2 - k4m_ * IO, t_C2_IC2 = k52 * C2 - k62 * IC2;
  const R t_C1_C2 = k1 * C1 - k2 * C2, t_C2_O = k31 * C2 - k41 * O, t_O_IO = k53 * O - k63 * IO;
  const R t_bO = bind * O - Ku * Obound, t_bIO = bind * IO - unbindIO * IObound;
  const R t_trO = trap * Cbound - Kt * Obound, t_trIO = trap * Cbound - Kt * IObound;
  f[34] = (double)(-t_IC1_IC2 + t_C1_IC1);
  f[35] = (double)(t_IC1_IC2 - t_IC2_IO + t_C2_IC2);
  f[36] = (double)(-t_C1_C2 - t_C1_IC1);
  f[37] = (double)(t_C1_C2 - t_C2_O - t_C2_IC2);
  f[38] = (double)(t_C2_O - t_O_IO - t_bO);
  f[39] = (double)(t_IC2_IO + t_O_IO - t_bIO);
  f[40] = (double)(t_bIO + t_trIO);
  f[41] = (double)(t_bO + t_trO);
  f[42] = (double)(-t_trO - t_trIO);
  f[43] = 0.0;

  // ---- gates (:1069-1098): (inf - state) formed in double
#define GRATE_DIV(i, inf, tau) (ga[i] = (double)Mx::rcp(tau), f[i] = ((double)(inf) - y[i]) * ga[i])
#define GRATE_MUL(i, inf, tauinv) (ga[i] = (double)(tauinv), f[i] = ((double)(inf) - y[i]) * ga[i])
  GRATE_MUL(17, hLss, c<R>(1 / thL));
  GRATE_MUL(18, hLssp, c<R>(1 / thLp));
  GRATE_MUL(10, mss, tm_inv);
  GRATE_MUL(11, hss, thf_inv);
  GRATE_MUL(12, hss, ths_inv);
  GRATE_DIV(19, ass, ta);
  GRATE_DIV(25, dss, td);
  GRATE_DIV(26, fss, tff);
  GRATE_DIV(27, fss, tfs);
  GRATE_MUL(30, fss, c<R>(1 / tjca));
  f[33] = (double)(anca * c<R>(k2n)) - y[33] * (double)jca;
  GRATE_DIV(44, xs1ss, txs1);
  GRATE_MUL(46, xk1ss, txk1_inv);
  f[1] = (double)(c<R>(aCaMK) * CaMKb * (CaMKb + camkt)) - bCaMK * y[1];
  GRATE_DIV(13, hss, tj);
  GRATE_DIV(22, assp, ta);
  GRATE_DIV(28, fss, tfcaf);
  GRATE_DIV(29, fss, tfcas);
  GRATE_DIV(31, fss, c<R>(2.5) * tff);
  GRATE_MUL(45, xs1ss, txs2_inv);
  GRATE_MUL(14, hssp, ths_inv * c<R>(1 / 3.0));
  GRATE_DIV(15, hss, c<R>(1.46) * tj);
  GRATE_MUL(16, mLss, tm_inv);
  GRATE_DIV(32, fss, c<R>(2.5) * tfcaf);
  GRATE_DIV(20, iss, tiF);
  GRATE_DIV(21, iss, tiS);
  GRATE_DIV(23, iss, dti * tiF);
  GRATE_DIV(24, iss, dti * tiS);
  GRATE_DIV(47, Jrel_inf, tau_rel);
  GRATE_DIV(48, Jrel_infp, tau_relp);
#undef GRATE_DIV
#undef GRATE_MUL

  // ---- concentrations and membrane potential in double (:1099-1107)
  const double dINa = INa, dINaL = INaL, dIto = Ito, dICaL = ICaL, dICaNa = ICaNa, dICaK = ICaK;
  const double dIKr = IKr, dIKs = IKs, dIK1 = IK1, dINaCa_i = INaCa_i, dINaCa_ss = INaCa_ss;
  const double dINaK = INaK, dINab = INab, dIKb = IKb, dIpCa = IpCa, dICab = ICab;
  const double CaMKa_d = CaMKa;
  const double fJp = 1.0 / (1.0 + KmCaMK / CaMKa_d);
  const double Jrel = (1.0 - fJp) * y[47] + fJp * y[48];
  const double cai_d = y[9], cass_d = y[2];
  const double Jupnp = (double)P(OP_UPSCALE) * 0.004375 * cai_d / (cai_d + 0.00092);
  const double Jupp = (double)P(OP_UPSCALE) * 2.75 * 0.004375 * cai_d / ((cai_d + 0.00092) - 0.00017);
  const double Jup = ((1.0 - fJp) * Jupnp + fJp * Jupp) - 0.0039375 / 15.0 * y[7];
  const double Jtr = (y[7] - y[8]) * 0.01;
  const double Jdiff = (cass_d - cai_d) * 5.0;
  const double JdiffNa = (y[4] - y[3]) * 0.5, JdiffK = (y[6] - y[5]) * 0.5;
  const double Bcass = 1.0 / (1.0 + BSRmax * KmBSR / ((KmBSR + cass_d) * (KmBSR + cass_d)) + BSLmax * KmBSL / ((KmBSL + cass_d) * (KmBSL + cass_d)));
  const double cmdnmax = P(OP_CMDNMAX);
  const double Bcai = 1.0 / (1.0 + cmdnmax * kmcmdn / ((kmcmdn + cai_d) * (kmcmdn + cai_d)) + trpnmax * kmtrpn / ((kmtrpn + cai_d) * (kmtrpn + cai_d)));
  const double Bcajsr = 1.0 / (1.0 + csqnmax * kmcsqn / ((kmcsqn + y[8]) * (kmcsqn + y[8])));
  constexpr double cA_Fmyo = cm * Acap / (Far * vmyo), cA_Fss = cm * Acap / (Far * vss);
  f[5] = -((dIto + dIKr + dIKs + dIK1 + dIKb + Istim) - 2.0 * dINaK) * cA_Fmyo + JdiffK * (vss / vmyo);
  f[6] = -dICaK * cA_Fss - JdiffK;
  f[3] = -(dINa + dINaL + 3.0 * dINaCa_i + 3.0 * dINaK + dINab) * cA_Fmyo + JdiffNa * (vss / vmyo);
  f[4] = -(dICaNa + 3.0 * dINaCa_ss) * cA_Fss - JdiffNa;
  f[0] = -(dINa + dINaL + dIto + dICaL + dICaNa + dICaK + dIKr + dIKs + dIK1 + dINaCa_i + dINaCa_ss +
           dINaK + dINab + dIKb + dIpCa + dICab + Istim);
  f[2] = Bcass * ((-(dICaL - 2.0 * dINaCa_ss) * (0.5 * cA_Fss) + Jrel * (vjsr / vss)) - Jdiff);
  f[9] = Bcai * ((-((dIpCa + dICab) - 2.0 * dINaCa_i) * (0.5 * cA_Fmyo) - Jup * (vnsr / vmyo)) + Jdiff * (vss / vmyo));
  f[7] = Jup - Jtr * (vjsr / vnsr);
  f[8] = Bcajsr * (Jtr - Jrel);
  f[49] = dINaL + dICaL + dIto + dIKr + dIKs + dIK1;
  f[50] = dINaL;
  f[51] = dICaL;
  if (cur) {
    cur->INa = dINa; cur->INaL = dINaL; cur->Ito = dIto; cur->ICaL = dICaL; cur->ICaNa = dICaNa;
    cur->ICaK = dICaK; cur->IKr = dIKr; cur->IKs = dIKs; cur->IK1 = dIK1; cur->INaCa_i = dINaCa_i;
    cur->INaCa_ss = dINaCa_ss; cur->INaK = dINaK; cur->INab = dINab; cur->IKb = dIKb;
    cur->IpCa = dIpCa; cur->ICab = dICab; cur->Istim = Istim; cur->Jrel = Jrel; cur->Jup = Jup;
    cur->CaMKa = CaMKa_d;
  }
}

// The 9x9 transition matrix of the Markov IKr block (states 34..42: IC1, IC2, C1, C2, O,
// IO, IObound, Obound, Cbound) at voltage v: d/dt S = A S, the fluxes of rhs() above
// (:1059-1068) written as a matrix (columns sum to zero).  The matrix is FP64; its
// voltage-dependent rate constants exp(B v) are evaluated in the rate precision R.
template <typename R = double, typename PA>
CEP_D void markov_matrix(double v, const PA &P, double (&A)[9][9])
{
  using Mx = M<R>;
  double k[14];
#pragma unroll
  for (int i = 0; i < 14; ++i) k[i] = g_markov.AQ[i] * (double)Mx::ex((R)(g_markov.B[i] * v));
  const double k1 = k[0], k2 = k[1], k3m = k[2], k4m = k[3], k11 = k[4], k21 = k[5], k31 = k[6], k41 = k[7];
  const double k51 = k[8], k52 = k[9], k53 = k[10], k61 = k[11], k62 = k[12], k63 = k[13];
  const double Ku = P(OP_KU), Kt = P(OP_KT), bind = P(OP_BIND);
  const double trap = Kt / (1.0 + exp(-(v - P(OP_VHALF)) * (1 / 6.789)));
  const double unbindIO = Ku * k53 / k63;
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = 0; j < 9; ++j) A[i][j] = 0.0;
  A[0][0] = -k11 - k61; A[0][1] = k21; A[0][2] = k51;
  A[1][0] = k11; A[1][1] = -k21 - k3m - k62; A[1][3] = k52; A[1][5] = k4m;
  A[2][0] = k61; A[2][2] = -k1 - k51; A[2][3] = k2;
  A[3][1] = k62; A[3][2] = k1; A[3][3] = -k2 - k31 - k52; A[3][4] = k41;
  A[4][3] = k31; A[4][4] = -k41 - k53 - bind; A[4][5] = k63; A[4][7] = Ku;
  A[5][1] = k3m; A[5][4] = k53; A[5][5] = -k4m - k63 - bind; A[5][6] = unbindIO;
  A[6][5] = bind; A[6][6] = -unbindIO - Kt; A[6][8] = trap;
  A[7][4] = bind; A[7][7] = -Ku - Kt; A[7][8] = trap;
  A[8][6] = Kt; A[8][7] = Kt; A[8][8] = -2.0 * trap;
}

}  // namespace ord
}  // namespace cep
