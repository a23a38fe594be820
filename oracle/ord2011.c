/*
 * ord2011.c -- ORACLE (test infrastructure only): FP64 restatement of the O'Hara-Rudy 2011
 * ventricular model as used by the DrugSimulation driver,
 *   DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp (state / constant / algebraic indices from
 *   DrugSimulation/enums/enum_Ohara_Rudy_2011.hpp),
 * together with its Rush-Larsen step (solveAnalytical, :991-1095), its adaptive step rule
 * (set_time_step, :1097-1137) and the driver loop with its biomarkers
 * (DrugSimulation/modules/drug_sim.cpp:306-420).
 *
 * Arrays use the reference's CellML layout: STATES[41], CONSTANTS[144], ALGEBRAIC[198]
 * (only the slots the step and the driver read are filled).  The reference's behaviour is
 * kept where it departs from the published model:
 *   - ___initConsts resets CONSTANTS[celltype] to 0 after initConsts(celltype) stored it, so
 *     every cell gets the endocardial constants (:496, :687-692);
 *   - ___applyDutta scales CONSTANTS at the index values of the ALGEBRAIC names IKs, IKr, IK1,
 *     ICaL, INaL (98, 96, 100, 81, 64 = Aff, GNaL, tjca, H, Gncx_b), after the derived
 *     constants were formed (:460-467);
 *   - the Rush-Larsen update of fcafp uses exp(-d / tfcafp) with d the STATES index of the
 *     d gate (25), not dt (:1064).
 * This copy of the model does not build here (its header pulls in curl).  The restatement is
 * pinned by the checks in tests/test_ord2011.py: golden vectors of the reference's patch-clamp
 * copy of the same model (CEPPatch_CiPA_Sim/patches/Ohara_Rudy_2011.cpp, compiled by
 * oracle/Makefile) for the constants, the Hill block and every rate and current that copy
 * completes; closed forms, the Rush-Larsen/rates consistency, the step rule, the currents this
 * model shares with the ORd-CiPA oracle (itself pinned by the compiled ORd-CiPA reference), and
 * the reference's logged EAD answers.
 */
#ifndef _POSIX_C_SOURCE
#define _POSIX_C_SOURCE 200809L
#endif
#include <math.h>
#include <pthread.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

/* CONSTANTS indices (enum E_CONSTANTS_T) */
enum {
  K_celltype = 0, K_nao, K_cao, K_ko, K_R, K_T, K_F, K_zna, K_zca, K_zk, K_L, K_rad, K_amp, K_duration,
  K_KmCaMK, K_aCaMK, K_bCaMK, K_CaMKo, K_KmCaM, K_cmdnmax_b, K_kmcmdn, K_trpnmax, K_kmtrpn, K_BSRmax,
  K_KmBSR, K_BSLmax, K_KmBSL, K_csqnmax, K_kmcsqn, K_cm, K_PKNa, K_mssV1, K_mssV2, K_mtV1, K_mtV2,
  K_mtD1, K_mtD2, K_mtV3, K_mtV4, K_hssV1, K_hssV2, K_Ahf, K_GNa, K_thL, K_GNaL_b, K_Gto_b, K_Kmn,
  K_k2n, K_PCa_b, K_GKr_b, K_GKs_b, K_GK1_b, K_kna1, K_kna2, K_kna3, K_kasymm, K_wna, K_wca, K_wnaca,
  K_kcaon, K_kcaoff, K_qna, K_qca, K_KmCaAct, K_Gncx_b, K_k1p, K_k1m, K_k2p, K_k2m, K_k3p, K_k3m,
  K_k4p, K_k4m, K_Knai0, K_Knao0, K_delta, K_Kki, K_Kko, K_MgADP, K_MgATP, K_Kmgatp, K_H, K_eP,
  K_Khp, K_Knap, K_Kxkur, K_Pnak_b, K_GKb_b, K_PNab, K_PCab, K_GpCa, K_KmCap, K_bt, K_cmdnmax, K_Ahs,
  K_thLp, K_GNaL, K_Gto, K_Aff, K_PCa, K_tjca, K_GKr, K_GKs, K_GK1, K_vcell, K_GKb, K_a_rel, K_btp,
  K_upScale, K_Afs, K_PCap, K_PCaNa, K_PCaK, K_Ageo, K_a_relp, K_PCaNap, K_PCaKp, K_Acap, K_vmyo,
  K_vnsr, K_vjsr, K_vss, K_h10_i, K_h11_i, K_h12_i, K_k1_i, K_k2_i, K_k5_i, K_Gncx, K_h10_ss,
  K_h11_ss, K_h12_ss, K_k1_ss, K_k2_ss, K_k5_ss, K_b1, K_a2, K_a4, K_Pnak, K_stim_start, K_BCL,
  K_cumm_BCL
};

void ord11_init_consts(double *C, double *S, double celltype)
{
  memset(C, 0, sizeof(double) * ORD11_NC);
  memset(S, 0, sizeof(double) * ORD11_NS);
  C[K_celltype] = celltype;
  /* ___initConsts (:494-679), which starts by setting celltype back to 0 */
  C[K_celltype] = 0;
  C[K_nao] = 140; C[K_cao] = 1.8; C[K_ko] = 5.4; C[K_R] = 8314; C[K_T] = 310; C[K_F] = 96485;
  C[K_zna] = 1; C[K_zca] = 2; C[K_zk] = 1; C[K_L] = 0.01; C[K_rad] = 0.0011;
  C[K_stim_start] = 10.0; C[K_BCL] = 1000.0; C[K_amp] = -80; C[K_duration] = 0.5;
  C[K_KmCaMK] = 0.15; C[K_aCaMK] = 0.05; C[K_bCaMK] = 0.00068; C[K_CaMKo] = 0.05; C[K_KmCaM] = 0.0015;
  C[K_cmdnmax_b] = 0.05; C[K_kmcmdn] = 0.00238; C[K_trpnmax] = 0.07; C[K_kmtrpn] = 0.0005;
  C[K_BSRmax] = 0.047; C[K_KmBSR] = 0.00087; C[K_BSLmax] = 1.124; C[K_KmBSL] = 0.0087;
  C[K_csqnmax] = 10; C[K_kmcsqn] = 0.8; C[K_cm] = 1; C[K_PKNa] = 0.01833;
  C[K_mssV1] = 39.57; C[K_mssV2] = 9.871; C[K_mtV1] = 11.64; C[K_mtV2] = 34.77; C[K_mtD1] = 6.765;
  C[K_mtD2] = 8.552; C[K_mtV3] = 77.42; C[K_mtV4] = 5.955; C[K_hssV1] = 82.9; C[K_hssV2] = 6.086;
  C[K_Ahf] = 0.99; C[K_GNa] = 75; C[K_thL] = 200; C[K_GNaL_b] = 0.0075; C[K_Gto_b] = 0.02;
  C[K_Kmn] = 0.002; C[K_k2n] = 1000; C[K_PCa_b] = 0.0001; C[K_GKr_b] = 0.046; C[K_GKs_b] = 0.0034;
  C[K_GK1_b] = 0.1908; C[K_kna1] = 15; C[K_kna2] = 5; C[K_kna3] = 88.12; C[K_kasymm] = 12.5;
  C[K_wna] = 6e4; C[K_wca] = 6e4; C[K_wnaca] = 5e3; C[K_kcaon] = 1.5e6; C[K_kcaoff] = 5e3;
  C[K_qna] = 0.5224; C[K_qca] = 0.167; C[K_KmCaAct] = 150e-6; C[K_Gncx_b] = 0.0008;
  C[K_k1p] = 949.5; C[K_k1m] = 182.4; C[K_k2p] = 687.2; C[K_k2m] = 39.4; C[K_k3p] = 1899;
  C[K_k3m] = 79300; C[K_k4p] = 639; C[K_k4m] = 40; C[K_Knai0] = 9.073; C[K_Knao0] = 27.78;
  C[K_delta] = -0.155; C[K_Kki] = 0.5; C[K_Kko] = 0.3582; C[K_MgADP] = 0.05; C[K_MgATP] = 9.8;
  C[K_Kmgatp] = 1.698e-7; C[K_H] = 1e-7; C[K_eP] = 4.2; C[K_Khp] = 1.698e-7; C[K_Knap] = 224;
  C[K_Kxkur] = 292; C[K_Pnak_b] = 30; C[K_GKb_b] = 0.003; C[K_PNab] = 3.75e-10; C[K_PCab] = 2.5e-8;
  C[K_GpCa] = 0.0005; C[K_KmCap] = 0.0005; C[K_bt] = 4.75;
  /* initial states */
  S[0] = -87; S[1] = 0; S[2] = 1e-4; S[3] = 7; S[4] = 7; S[5] = 145; S[6] = 145; S[7] = 1.2;
  S[8] = 1.2; S[9] = 1e-4; S[10] = 0; S[11] = 1; S[12] = 1; S[13] = 1; S[14] = 1; S[15] = 1;
  S[16] = 0; S[17] = 1; S[18] = 1; S[19] = 0; S[20] = 1; S[21] = 1; S[22] = 0; S[23] = 1;
  S[24] = 1; S[25] = 0; S[26] = 1; S[27] = 1; S[28] = 1; S[29] = 1; S[30] = 1; S[31] = 1;
  S[32] = 1; S[33] = 0; S[34] = 0; S[35] = 0; S[36] = 0; S[37] = 0; S[38] = 1; S[39] = 0; S[40] = 0;
  /* derived constants (celltype-dependent branches evaluate with celltype == 0) */
  const double ct = C[K_celltype];
  C[K_cmdnmax] = (ct == 1.0) ? C[K_cmdnmax_b] * 1.3 : C[K_cmdnmax_b];
  C[K_Ahs] = 1.0 - C[K_Ahf];
  C[K_thLp] = 3.0 * C[K_thL];
  C[K_GNaL] = (ct == 1.0) ? C[K_GNaL_b] * 0.6 : C[K_GNaL_b];
  C[K_Gto] = (ct == 1.0) ? C[K_Gto_b] * 4.0 : (ct == 2.0) ? C[K_Gto_b] * 4.0 : C[K_Gto_b];
  C[K_Aff] = 0.6;
  C[K_PCa] = (ct == 1.0) ? C[K_PCa_b] * 1.2 : (ct == 2.0) ? C[K_PCa_b] * 2.5 : C[K_PCa_b];
  C[K_tjca] = 75.0;
  C[K_GKr] = (ct == 1.0) ? C[K_GKr_b] * 1.3 : (ct == 2.0) ? C[K_GKr_b] * 0.8 : C[K_GKr_b];
  C[K_GKs] = (ct == 1.0) ? C[K_GKs_b] * 1.4 : C[K_GKs_b];
  C[K_GK1] = (ct == 1.0) ? C[K_GK1_b] * 1.2 : (ct == 2.0) ? C[K_GK1_b] * 1.3 : C[K_GK1_b];
  C[K_vcell] = 1000.0 * 3.14 * C[K_rad] * C[K_rad] * C[K_L];
  C[K_GKb] = (ct == 1.0) ? C[K_GKb_b] * 0.6 : C[K_GKb_b];
  C[K_a_rel] = 0.5 * C[K_bt];
  C[K_btp] = 1.25 * C[K_bt];
  C[K_upScale] = (ct == 1.0) ? 1.3 : 1.0;
  C[K_Afs] = 1.0 - C[K_Aff];
  C[K_PCap] = 1.1 * C[K_PCa];
  C[K_PCaNa] = 0.00125 * C[K_PCa];
  C[K_PCaK] = 0.0003574 * C[K_PCa];
  C[K_Ageo] = 2.0 * 3.14 * C[K_rad] * C[K_rad] + 2.0 * 3.14 * C[K_rad] * C[K_L];
  C[K_a_relp] = 0.5 * C[K_btp];
  C[K_PCaNap] = 0.00125 * C[K_PCap];
  C[K_PCaKp] = 0.0003574 * C[K_PCap];
  C[K_Acap] = 2.0 * C[K_Ageo];
  C[K_vmyo] = 0.68 * C[K_vcell];
  C[K_vnsr] = 0.0552 * C[K_vcell];
  C[K_vjsr] = 0.0048 * C[K_vcell];
  C[K_vss] = 0.02 * C[K_vcell];
  C[K_h10_i] = C[K_kasymm] + 1.0 + (C[K_nao] / C[K_kna1]) * (1.0 + C[K_nao] / C[K_kna2]);
  C[K_h11_i] = (C[K_nao] * C[K_nao]) / (C[K_h10_i] * C[K_kna1] * C[K_kna2]);
  C[K_h12_i] = 1.0 / C[K_h10_i];
  C[K_k1_i] = C[K_h12_i] * C[K_cao] * C[K_kcaon];
  C[K_k2_i] = C[K_kcaoff];
  C[K_k5_i] = C[K_kcaoff];
  C[K_Gncx] = (ct == 1.0) ? C[K_Gncx_b] * 1.1 : (ct == 2.0) ? C[K_Gncx_b] * 1.4 : C[K_Gncx_b];
  C[K_h10_ss] = C[K_kasymm] + 1.0 + (C[K_nao] / C[K_kna1]) * (1.0 + C[K_nao] / C[K_kna2]);
  C[K_h11_ss] = (C[K_nao] * C[K_nao]) / (C[K_h10_ss] * C[K_kna1] * C[K_kna2]);
  C[K_h12_ss] = 1.0 / C[K_h10_ss];
  C[K_k1_ss] = C[K_h12_ss] * C[K_cao] * C[K_kcaon];
  C[K_k2_ss] = C[K_kcaoff];
  C[K_k5_ss] = C[K_kcaoff];
  C[K_b1] = C[K_k1m] * C[K_MgADP];
  C[K_a2] = C[K_k2p];
  C[K_a4] = ((C[K_k4p] * C[K_MgATP]) / C[K_Kmgatp]) / (1.0 + C[K_MgATP] / C[K_Kmgatp]);
  C[K_Pnak] = (ct == 1.0) ? C[K_Pnak_b] * 0.9 : (ct == 2.0) ? C[K_Pnak_b] * 0.7 : C[K_Pnak_b];
}

/* ___applyDutta (:460-467): the scaled slots are the ALGEBRAIC indices of IKs, IKr, IK1, ICaL,
   INaL used on CONSTANTS */
void ord11_apply_dutta(double *C)
{
  C[98] *= 1.870;
  C[96] *= 1.013;
  C[100] *= 1.698;
  C[81] *= 1.007;
  C[64] *= 2.661;
}

/* ___applyDrugEffect (:469-485): Hill block of PCa, GK1, GKs, GNa, GNaL, Gto, GKr from the
   IC50 CSV row order (IC50, h) x 7, skipped when either value is <= epsilon = 10E-14 */
void ord11_apply_drug(double *C, double conc, const double *ic50)
{
  static const int slot[7] = { K_PCa, K_GK1, K_GKs, K_GNa, K_GNaL, K_Gto, K_GKr };
  const double eps = 10E-14;
  for (int k = 0; k < 7; ++k) {
    const double i50 = ic50[2 * k], h = ic50[2 * k + 1];
    if (i50 > eps && h > eps) C[slot[k]] *= 1.0 / (1.0 + pow(conc / i50, h));
  }
}

/* ALGEBRAIC slots (enum E_ALGEBRAIC_T) written by ord11_rates */
enum {
  A_mss = 0, A_hss = 1, A_hLss = 2, A_hLssp = 3, A_ass = 4, A_iss = 5, A_dss = 6, A_fss = 7,
  A_km2n = 8, A_xrss = 9, A_xs1ss = 10, A_xk1ss = 11, A_Istim = 12, A_tm = 13, A_thf = 14,
  A_ths = 15, A_jss = 16, A_ta = 17, A_fcass = 19, A_anca = 20, A_td = 21, A_tff = 22, A_tfs = 23,
  A_txrf = 24, A_txrs = 25, A_xs2ss = 26, A_txs1 = 27, A_txk1 = 28, A_tj = 30, A_hssp = 31,
  A_mLss = 32, A_assp = 34, A_tfcaf = 35, A_tfcas = 36, A_tffp = 37, A_txs2 = 38, A_thsp = 40,
  A_tjp = 41, A_tmL = 42, A_tfcafp = 44, A_tiF = 46, A_tiS = 48, A_tiFp = 54, A_tiSp = 55,
  A_INa = 62, A_INaL = 64, A_Ito = 70, A_ICaL = 81, A_ICaNa = 82, A_ICaK = 85, A_Jrel_inf = 86,
  A_Jrel_infp = 87, A_tau_rel = 92, A_tau_relp = 93, A_IKr = 96, A_IKs = 98, A_IK1 = 100,
  A_INaCa_i = 132, A_INaCa_ss = 162, A_INaK = 181, A_IKb = 183, A_INab = 184, A_ICab = 186,
  A_IpCa = 188, A_Jrel = 191, A_Jup = 196
};

/* one Na/Ca exchanger compartment (:870-912 for _i, :921-961 for _ss) */
static double ncx(const double *C, double na, double ca, double hna, double hca, double h11,
                  double k1, double k2, double k5, double allo, double frac)
{
  const double h7 = 1.0 + (C[K_nao] / C[K_kna3]) * (1.0 + 1.0 / hna);
  const double h8 = C[K_nao] / (C[K_kna3] * hna * h7);
  const double k3pp = h8 * C[K_wnaca];
  const double h1 = 1.0 + (na / C[K_kna3]) * (1.0 + hna);
  const double h2 = (na * hna) / (C[K_kna3] * h1);
  const double k4pp = h2 * C[K_wnaca];
  const double h4 = 1.0 + (na / C[K_kna1]) * (1.0 + na / C[K_kna2]);
  const double h5 = (na * na) / (h4 * C[K_kna1] * C[K_kna2]);
  const double k7 = h5 * h2 * C[K_wna];
  const double k8 = h8 * h11 * C[K_wna];
  const double h9 = 1.0 / h7;
  const double k3p = h9 * C[K_wca];
  const double k3 = k3p + k3pp;
  const double h3 = 1.0 / h1;
  const double k4p = (h3 * C[K_wca]) / hca;
  const double k4 = k4p + k4pp;
  const double h6 = 1.0 / h4;
  const double k6 = h6 * ca * C[K_kcaon];
  const double x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
  const double x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
  const double x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
  const double x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
  const double E1 = x1 / (x1 + x2 + x3 + x4);
  const double E2 = x2 / (x1 + x2 + x3 + x4);
  const double E3 = x3 / (x1 + x2 + x3 + x4);
  const double E4 = x4 / (x1 + x2 + x3 + x4);
  const double JNa = (3.0 * (E4 * k7 - E1 * k8) + E3 * k4pp) - E2 * k3pp;
  const double JCa = E2 * k2 - E1 * k1;
  return frac * C[K_Gncx] * allo * (C[K_zna] * JNa + C[K_zca] * JCa);
}

/* computeRates (:742-989), SINGLE_CELL stimulus; tissue_stim >= 0 selects the TISSUE branch
   (amp when nonzero, else 0) */
void ord11_rates_ex(double t, const double *C, double *R, const double *S, double *A, int tissue_stim)
{
  const double V = S[0], CaMKt = S[1], cass = S[2], nai = S[3], nass = S[4], ki = S[5], kss = S[6];
  const double cansr = S[7], cajsr = S[8], cai = S[9];
  const double m = S[10], hf = S[11], hs = S[12], j = S[13], hsp = S[14], jp = S[15], mL = S[16];
  const double hL = S[17], hLp = S[18], a = S[19], iF = S[20], iS = S[21], ap = S[22], iFp = S[23];
  const double iSp = S[24], d = S[25], ff = S[26], fs = S[27], fcaf = S[28], fcas = S[29];
  const double jca = S[30], ffp = S[31], fcafp = S[32], nca = S[33], xrf = S[34], xrs = S[35];
  const double xs1 = S[36], xs2 = S[37], xk1 = S[38], Jrelnp = S[39], Jrelp = S[40];
  const double RT_F = (C[K_R] * C[K_T]) / C[K_F];

  double Istim;
  if (tissue_stim >= 0) {
    Istim = tissue_stim ? C[K_amp] : 0.0;
  } else {
    const double ts = t - C[K_stim_start];
    Istim = (t >= C[K_stim_start] && ts - floor(ts / C[K_BCL]) * C[K_BCL] <= C[K_duration]) ? C[K_amp] : 0.0;
  }
  A[A_Istim] = Istim;
  /* gates: steady states and time constants */
  const double hLss = 1.0 / (1.0 + exp((V + 87.61) / 7.488));
  const double hLssp = 1.0 / (1.0 + exp((V + 93.81) / 7.488));
  const double mss = 1.0 / (1.0 + exp(-(V + C[K_mssV1]) / C[K_mssV2]));
  const double tm = 1.0 / (C[K_mtD1] * exp((V + C[K_mtV1]) / C[K_mtV2]) + C[K_mtD2] * exp(-(V + C[K_mtV3]) / C[K_mtV4]));
  const double hss = 1.0 / (1.0 + exp((V + C[K_hssV1]) / C[K_hssV2]));
  const double thf = 1.0 / (1.432e-05 * exp(-(V + 1.196) / 6.285) + 6.149 * exp((V + 0.5096) / 20.27));
  const double ths = 1.0 / (0.009794 * exp(-(V + 17.95) / 28.05) + 0.3343 * exp((V + 5.73) / 56.66));
  const double ass = 1.0 / (1.0 + exp(-(V - 14.34) / 14.82));
  const double ta = 1.0515 / (1.0 / (1.2089 * (1.0 + exp(-(V - 18.4099) / 29.3814))) + 3.5 / (1.0 + exp((V + 100.0) / 29.3814)));
  const double dss = 1.0 / (1.0 + exp(-(V + 3.94) / 4.23));
  const double td = 0.6 + 1.0 / (exp(-0.05 * (V + 6.0)) + exp(0.09 * (V + 14.0)));
  const double fss = 1.0 / (1.0 + exp((V + 19.58) / 3.696));
  const double tff = 7.0 + 1.0 / (0.0045 * exp(-(V + 20.0) / 10.0) + 0.0045 * exp((V + 20.0) / 10.0));
  const double tfs = 1000.0 + 1.0 / (3.5e-05 * exp(-(V + 5.0) / 4.0) + 3.5e-05 * exp((V + 5.0) / 6.0));
  const double fcass = fss;
  const double km2n = jca * 1.0;
  const double anca = 1.0 / (C[K_k2n] / km2n + pow(1.0 + C[K_Kmn] / cass, 4.0));
  const double xrss = 1.0 / (1.0 + exp(-(V + 8.337) / 6.789));
  const double txrf = 12.98 + 1.0 / (0.3652 * exp((V - 31.66) / 3.869) + 4.123e-05 * exp(-(V - 47.78) / 20.38));
  const double txrs = 1.865 + 1.0 / (0.06629 * exp((V - 34.70) / 7.355) + 1.128e-05 * exp(-(V - 29.74) / 25.94));
  const double xs1ss = 1.0 / (1.0 + exp(-(V + 11.60) / 8.932));
  const double txs1 = 817.3 + 1.0 / (2.326e-4 * exp((V + 48.28) / 17.80) + 0.001292 * exp(-(V + 210.0) / 230.0));
  const double xk1ss = 1.0 / (1.0 + exp(-(V + 2.5538 * C[K_ko] + 144.59) / (1.5692 * C[K_ko] + 3.8115)));
  const double txk1 = 122.2 / (exp(-(V + 127.2) / 20.36) + exp((V + 236.8) / 69.33));
  const double jss = hss;
  const double tj = 2.038 + 1.0 / (0.02136 * exp(-(V + 100.6) / 8.281) + 0.3052 * exp((V + 0.9941) / 38.45));
  const double assp = 1.0 / (1.0 + exp(-(V - 24.34) / 14.82));
  const double tfcaf = 7.0 + 1.0 / (0.04 * exp(-(V - 4.0) / 7.0) + 0.04 * exp((V - 4.0) / 7.0));
  const double tfcas = 100.0 + 1.0 / (0.00012 * exp(-V / 3.0) + 0.00012 * exp(V / 7.0));
  const double tffp = 2.5 * tff;
  const double xs2ss = xs1ss;
  const double txs2 = 1.0 / (0.01 * exp((V - 50.0) / 20.0) + 0.0193 * exp(-(V + 66.54) / 31.0));
  const double CaMKb = (C[K_CaMKo] * (1.0 - CaMKt)) / (1.0 + C[K_KmCaM] / cass);
  const double hssp = 1.0 / (1.0 + exp((V + 89.1) / 6.086));
  const double thsp = 3.0 * ths;
  const double tjp = 1.46 * tj;
  const double mLss = 1.0 / (1.0 + exp(-(V + 42.85) / 5.264));
  const double tmL = tm;
  const double tfcafp = 2.5 * tfcaf;
  const double iss = 1.0 / (1.0 + exp((V + 43.94) / 5.711));
  const double delta_epi = (C[K_celltype] == 1.0) ? 1.0 - 0.95 / (1.0 + exp((V + 70.0) / 5.0)) : 1.0;
  const double tiF = (4.562 + 1.0 / (0.3933 * exp(-(V + 100.0) / 100.0) + 0.08004 * exp((V + 50.0) / 16.59))) * delta_epi;
  const double tiS = (23.62 + 1.0 / (0.001416 * exp(-(V + 96.52) / 59.05) + 1.780e-8 * exp((V + 114.1) / 8.079))) * delta_epi;
  const double dti_develop = 1.354 + 1.0e-4 / (exp((V - 167.4) / 15.89) + exp(-(V - 12.23) / 0.2154));
  const double dti_recover = 1.0 - 0.5 / (1.0 + exp((V + 70.0) / 20.0));
  const double tiFp = dti_develop * dti_recover * tiF;
  const double tiSp = dti_develop * dti_recover * tiS;

  /* ICaL, ICaNa, ICaK (GHK driving forces as written, no 0/0 guard) */
  const double f = C[K_Aff] * ff + C[K_Afs] * fs;
  const double Afcaf = 0.3 + 0.6 / (1.0 + exp((V - 10.0) / 10.0));
  const double Afcas = 1.0 - Afcaf;
  const double fca = Afcaf * fcaf + Afcas * fcas;
  const double fp = C[K_Aff] * ffp + C[K_Afs] * fs;
  const double fcap = Afcaf * fcafp + Afcas * fcas;
  const double vffrt = (V * C[K_F] * C[K_F]) / (C[K_R] * C[K_T]);
  const double vfrt = (V * C[K_F]) / (C[K_R] * C[K_T]);
  const double PhiCaL = (4.0 * vffrt * (cass * exp(2.0 * vfrt) - 0.341 * C[K_cao])) / (exp(2.0 * vfrt) - 1.0);
  const double CaMKa = CaMKb + CaMKt;
  const double fICaLp = 1.0 / (1.0 + C[K_KmCaMK] / CaMKa);
  const double gnp = d * (f * (1.0 - nca) + jca * fca * nca);
  const double gp = d * (fp * (1.0 - nca) + jca * fcap * nca);
  const double ICaL = (1.0 - fICaLp) * C[K_PCa] * PhiCaL * gnp + fICaLp * C[K_PCap] * PhiCaL * gp;
  const double Jrel_inf_temp = (C[K_a_rel] * -ICaL) / (1.0 + 1.0 * pow(1.5 / cajsr, 8.0));
  const double Jrel_inf = (C[K_celltype] == 2.0) ? Jrel_inf_temp * 1.7 : Jrel_inf_temp;
  const double tau_rel_temp = C[K_bt] / (1.0 + 0.0123 / cajsr);
  const double tau_rel = (tau_rel_temp < 0.001) ? 0.001 : tau_rel_temp;
  const double Jrel_temp = (C[K_a_relp] * -ICaL) / (1.0 + pow(1.5 / cajsr, 8.0));
  const double Jrel_infp = (C[K_celltype] == 2.0) ? Jrel_temp * 1.7 : Jrel_temp;
  const double tau_relp_temp = C[K_btp] / (1.0 + 0.0123 / cajsr);
  const double tau_relp = (tau_relp_temp < 0.001) ? 0.001 : tau_relp_temp;
  /* K currents */
  const double EK = RT_F * log(C[K_ko] / ki);
  const double AiF = 1.0 / (1.0 + exp((V - 213.6) / 151.2));
  const double AiS = 1.0 - AiF;
  const double ii = AiF * iF + AiS * iS;
  const double ip = AiF * iFp + AiS * iSp;
  const double fItop = 1.0 / (1.0 + C[K_KmCaMK] / CaMKa);
  const double Ito = C[K_Gto] * (V - EK) * ((1.0 - fItop) * a * ii + fItop * ap * ip);
  const double Axrf = 1.0 / (1.0 + exp((V + 54.81) / 38.21));
  const double Axrs = 1.0 - Axrf;
  const double xr = Axrf * xrf + Axrs * xrs;
  const double rkr = ((1.0 / (1.0 + exp((V + 55.0) / 75.0))) * 1.0) / (1.0 + exp((V - 10.0) / 30.0));
  const double IKr = C[K_GKr] * pow(C[K_ko] / 5.4, 1.0 / 2) * xr * rkr * (V - EK);
  const double EKs = RT_F * log((C[K_ko] + C[K_PKNa] * C[K_nao]) / (ki + C[K_PKNa] * nai));
  const double KsCa = 1.0 + 0.6 / (1.0 + pow(3.8e-05 / cai, 1.4));
  const double IKs = C[K_GKs] * KsCa * xs1 * xs2 * (V - EKs);
  const double rk1 = 1.0 / (1.0 + exp(((V + 105.8) - 2.6 * C[K_ko]) / 9.493));
  const double IK1 = C[K_GK1] * pow(C[K_ko], 1.0 / 2) * rk1 * xk1 * (V - EK);
  /* INaK */
  const double Knao = C[K_Knao0] * exp(((1.0 - C[K_delta]) * V * C[K_F]) / (3.0 * C[K_R] * C[K_T]));
  const double a3 = (C[K_k3p] * pow(C[K_ko] / C[K_Kko], 2.0)) /
                    ((pow(1.0 + C[K_nao] / Knao, 3.0) + pow(1.0 + C[K_ko] / C[K_Kko], 2.0)) - 1.0);
  const double P = C[K_eP] / (1.0 + C[K_H] / C[K_Khp] + nai / C[K_Knap] + ki / C[K_Kxkur]);
  const double b3 = (C[K_k3m] * P * C[K_H]) / (1.0 + C[K_MgATP] / C[K_Kmgatp]);
  const double Knai = C[K_Knai0] * exp((C[K_delta] * V * C[K_F]) / (3.0 * C[K_R] * C[K_T]));
  const double a1 = (C[K_k1p] * pow(nai / Knai, 3.0)) / ((pow(1.0 + nai / Knai, 3.0) + pow(1.0 + ki / C[K_Kki], 2.0)) - 1.0);
  const double b2 = (C[K_k2m] * pow(C[K_nao] / Knao, 3.0)) /
                    ((pow(1.0 + C[K_nao] / Knao, 3.0) + pow(1.0 + C[K_ko] / C[K_Kko], 2.0)) - 1.0);
  const double b4 = (C[K_k4m] * pow(ki / C[K_Kki], 2.0)) / ((pow(1.0 + nai / Knai, 3.0) + pow(1.0 + ki / C[K_Kki], 2.0)) - 1.0);
  const double A2 = C[K_a2], A4 = C[K_a4], B1 = C[K_b1];
  const double x1 = A4 * a1 * A2 + b2 * b4 * b3 + A2 * b4 * b3 + b3 * a1 * A2;
  const double x2 = b2 * B1 * b4 + a1 * A2 * a3 + a3 * B1 * b4 + A2 * a3 * b4;
  const double x3 = A2 * a3 * A4 + b3 * b2 * B1 + b2 * B1 * A4 + a3 * A4 * B1;
  const double x4 = b4 * b3 * b2 + a3 * A4 * a1 + b2 * A4 * a1 + b3 * b2 * a1;
  const double E1 = x1 / (x1 + x2 + x3 + x4);
  const double E2 = x2 / (x1 + x2 + x3 + x4);
  const double JnakNa = 3.0 * (E1 * a3 - E2 * b3);
  const double E3 = x3 / (x1 + x2 + x3 + x4);
  const double E4 = x4 / (x1 + x2 + x3 + x4);
  const double JnakK = 2.0 * (E4 * B1 - E3 * a1);
  const double INaK = C[K_Pnak] * (C[K_zna] * JnakNa + C[K_zk] * JnakK);
  const double xkb = 1.0 / (1.0 + exp(-(V - 14.48) / 18.34));
  const double IKb = C[K_GKb] * xkb * (V - EK);
  const double JdiffK = (kss - ki) / 2.0;
  const double PhiCaK = (1.0 * vffrt * (0.75 * kss * exp(1.0 * vfrt) - 0.75 * C[K_ko])) / (exp(1.0 * vfrt) - 1.0);
  const double ICaK = (1.0 - fICaLp) * C[K_PCaK] * PhiCaK * gnp + fICaLp * C[K_PCaKp] * PhiCaK * gp;
  /* Na currents */
  const double ENa = RT_F * log(C[K_nao] / nai);
  const double h = C[K_Ahf] * hf + C[K_Ahs] * hs;
  const double hp = C[K_Ahf] * hf + C[K_Ahs] * hsp;
  const double fINap = 1.0 / (1.0 + C[K_KmCaMK] / CaMKa);
  const double INa = C[K_GNa] * (V - ENa) * pow(m, 3.0) * ((1.0 - fINap) * h * j + fINap * hp * jp);
  const double fINaLp = 1.0 / (1.0 + C[K_KmCaMK] / CaMKa);
  const double INaL = C[K_GNaL] * (V - ENa) * mL * ((1.0 - fINaLp) * hL + fINaLp * hLp);
  /* Na/Ca exchange */
  const double hna = exp((C[K_qna] * V * C[K_F]) / (C[K_R] * C[K_T]));
  const double hca = exp((C[K_qca] * V * C[K_F]) / (C[K_R] * C[K_T]));
  const double allo_i = 1.0 / (1.0 + pow(C[K_KmCaAct] / cai, 2.0));
  const double INaCa_i = ncx(C, nai, cai, hna, hca, C[K_h11_i], C[K_k1_i], C[K_k2_i], C[K_k5_i], allo_i, 0.8);
  const double INab = (C[K_PNab] * vffrt * (nai * exp(vfrt) - C[K_nao])) / (exp(vfrt) - 1.0);
  const double JdiffNa = (nass - nai) / 2.0;
  const double PhiCaNa = (1.0 * vffrt * (0.75 * nass * exp(1.0 * vfrt) - 0.75 * C[K_nao])) / (exp(1.0 * vfrt) - 1.0);
  const double ICaNa = (1.0 - fICaLp) * C[K_PCaNa] * PhiCaNa * gnp + fICaLp * C[K_PCaNap] * PhiCaNa * gp;
  const double allo_ss = 1.0 / (1.0 + pow(C[K_KmCaAct] / cass, 2.0));
  const double INaCa_ss = ncx(C, nass, cass, hna, hca, C[K_h11_ss], C[K_k1_ss], C[K_k2_ss], C[K_k5_ss], allo_ss, 0.2);
  const double IpCa = (C[K_GpCa] * cai) / (C[K_KmCap] + cai);
  const double ICab = (C[K_PCab] * 4.0 * vffrt * (cai * exp(2.0 * vfrt) - 0.341 * C[K_cao])) / (exp(2.0 * vfrt) - 1.0);
  const double Jdiff = (cass - cai) / 0.2;
  const double fJrelp = 1.0 / (1.0 + C[K_KmCaMK] / CaMKa);
  const double Jrel = (1.0 - fJrelp) * Jrelnp + fJrelp * Jrelp;
  const double Bcass = 1.0 / (1.0 + (C[K_BSRmax] * C[K_KmBSR]) / pow(C[K_KmBSR] + cass, 2.0) +
                              (C[K_BSLmax] * C[K_KmBSL]) / pow(C[K_KmBSL] + cass, 2.0));
  const double Jupnp = (C[K_upScale] * 0.004375 * cai) / (cai + 0.00092);
  const double Jupp = (C[K_upScale] * 2.75 * 0.004375 * cai) / ((cai + 0.00092) - 0.00017);
  const double fJupp = 1.0 / (1.0 + C[K_KmCaMK] / CaMKa);
  const double Jleak = (0.0039375 * cansr) / 15.0;
  const double Jup = ((1.0 - fJupp) * Jupnp + fJupp * Jupp) - Jleak;
  const double Bcai = 1.0 / (1.0 + (C[K_cmdnmax] * C[K_kmcmdn]) / pow(C[K_kmcmdn] + cai, 2.0) +
                             (C[K_trpnmax] * C[K_kmtrpn]) / pow(C[K_kmtrpn] + cai, 2.0));
  const double Jtr = (cansr - cajsr) / 100.0;
  const double Bcajsr = 1.0 / (1.0 + (C[K_csqnmax] * C[K_kmcsqn]) / pow(C[K_kmcsqn] + cajsr, 2.0));

  /* RATES (:961-988) */
  R[17] = (hLss - hL) / C[K_thL];
  R[18] = (hLssp - hLp) / C[K_thLp];
  R[10] = (mss - m) / tm;
  R[11] = (hss - hf) / thf;
  R[12] = (hss - hs) / ths;
  R[19] = (ass - a) / ta;
  R[25] = (dss - d) / td;
  R[26] = (fss - ff) / tff;
  R[27] = (fss - fs) / tfs;
  R[30] = (fcass - jca) / C[K_tjca];
  R[33] = anca * C[K_k2n] - nca * km2n;
  R[34] = (xrss - xrf) / txrf;
  R[35] = (xrss - xrs) / txrs;
  R[36] = (xs1ss - xs1) / txs1;
  R[38] = (xk1ss - xk1) / txk1;
  R[13] = (jss - j) / tj;
  R[22] = (assp - ap) / ta;
  R[28] = (fcass - fcaf) / tfcaf;
  R[29] = (fcass - fcas) / tfcas;
  R[31] = (fss - ffp) / tffp;
  R[37] = (xs2ss - xs2) / txs2;
  R[1] = C[K_aCaMK] * CaMKb * (CaMKb + CaMKt) - C[K_bCaMK] * CaMKt;
  R[14] = (hssp - hsp) / thsp;
  R[15] = (jss - jp) / tjp;
  R[16] = (mLss - mL) / tmL;
  R[32] = (fcass - fcafp) / tfcafp;
  R[20] = (iss - iF) / tiF;
  R[21] = (iss - iS) / tiS;
  R[23] = (iss - iFp) / tiFp;
  R[24] = (iss - iSp) / tiSp;
  R[39] = (Jrel_inf - Jrelnp) / tau_rel;
  R[40] = (Jrel_infp - Jrelp) / tau_relp;
  R[5] = (-((Ito + IKr + IKs + IK1 + IKb + Istim) - 2.0 * INaK) * C[K_cm] * C[K_Acap]) / (C[K_F] * C[K_vmyo]) +
         (JdiffK * C[K_vss]) / C[K_vmyo];
  R[6] = (-ICaK * C[K_cm] * C[K_Acap]) / (C[K_F] * C[K_vss]) - JdiffK;
  R[3] = (-(INa + INaL + 3.0 * INaCa_i + 3.0 * INaK + INab) * C[K_Acap] * C[K_cm]) / (C[K_F] * C[K_vmyo]) +
         (JdiffNa * C[K_vss]) / C[K_vmyo];
  R[4] = (-(ICaNa + 3.0 * INaCa_ss) * C[K_cm] * C[K_Acap]) / (C[K_F] * C[K_vss]) - JdiffNa;
  R[0] = -(INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa_i + INaCa_ss + INaK + INab + IKb +
           IpCa + ICab + Istim);
  R[2] = Bcass * (((-(ICaL - 2.0 * INaCa_ss) * C[K_cm] * C[K_Acap]) / (2.0 * C[K_F] * C[K_vss]) +
                   (Jrel * C[K_vjsr]) / C[K_vss]) - Jdiff);
  R[9] = Bcai * (((-((IpCa + ICab) - 2.0 * INaCa_i) * C[K_cm] * C[K_Acap]) / (2.0 * C[K_F] * C[K_vmyo]) -
                  (Jup * C[K_vnsr]) / C[K_vmyo]) + (Jdiff * C[K_vss]) / C[K_vmyo]);
  R[7] = Jup - (Jtr * C[K_vjsr]) / C[K_vnsr];
  R[8] = Bcajsr * (Jtr - Jrel);

  /* ALGEBRAIC slots used by the step and the drivers */
  A[A_mss] = mss; A[A_tm] = tm; A[A_hss] = hss; A[A_thf] = thf; A[A_ths] = ths; A[A_jss] = jss;
  A[A_tj] = tj; A[A_hssp] = hssp; A[A_thsp] = thsp; A[A_tjp] = tjp; A[A_mLss] = mLss; A[A_tmL] = tmL;
  A[A_hLss] = hLss; A[A_hLssp] = hLssp; A[A_ass] = ass; A[A_ta] = ta; A[A_iss] = iss; A[A_tiF] = tiF;
  A[A_tiS] = tiS; A[A_assp] = assp; A[A_tiFp] = tiFp; A[A_tiSp] = tiSp; A[A_dss] = dss; A[A_td] = td;
  A[A_fss] = fss; A[A_tff] = tff; A[A_tfs] = tfs; A[A_fcass] = fcass; A[A_tfcaf] = tfcaf;
  A[A_tfcas] = tfcas; A[A_tffp] = tffp; A[A_tfcafp] = tfcafp; A[A_anca] = anca; A[A_km2n] = km2n;
  A[A_xrss] = xrss; A[A_txrf] = txrf; A[A_txrs] = txrs; A[A_xs1ss] = xs1ss; A[A_txs1] = txs1;
  A[A_xs2ss] = xs2ss; A[A_txs2] = txs2; A[A_xk1ss] = xk1ss; A[A_txk1] = txk1;
  A[A_Jrel_inf] = Jrel_inf; A[A_tau_rel] = tau_rel; A[A_Jrel_infp] = Jrel_infp; A[A_tau_relp] = tau_relp;
  A[A_INa] = INa; A[A_INaL] = INaL; A[A_Ito] = Ito; A[A_ICaL] = ICaL; A[A_ICaNa] = ICaNa; A[A_ICaK] = ICaK;
  A[A_IKr] = IKr; A[A_IKs] = IKs; A[A_IK1] = IK1; A[A_INaCa_i] = INaCa_i; A[A_INaCa_ss] = INaCa_ss;
  A[A_INaK] = INaK; A[A_IKb] = IKb; A[A_INab] = INab; A[A_ICab] = ICab; A[A_IpCa] = IpCa;
  A[A_Jrel] = Jrel; A[A_Jup] = Jup;
}

void ord11_rates(double t, const double *C, double *R, const double *S, double *A)
{
  ord11_rates_ex(t, C, R, S, A, -1);
}

/* solveAnalytical (:991-1095, non-EULER branch): Rush-Larsen for the gates and the two Jrel
   states with the ALGEBRAIC values of the last ord11_rates call, forward Euler for CaMKt, V
   and the concentrations with its RATES */
void ord11_rush_larsen(double dt, const double *C, const double *R, double *S, const double *A)
{
#define RL(i, inf, tau) S[i] = A[inf] - (A[inf] - S[i]) * exp(-dt / A[tau])
  RL(10, A_mss, A_tm);
  RL(11, A_hss, A_thf);
  RL(12, A_hss, A_ths);
  RL(13, A_jss, A_tj);
  RL(14, A_hssp, A_thsp);
  RL(15, A_jss, A_tjp);
  RL(16, A_mLss, A_tmL);
  S[17] = A[A_hLss] - (A[A_hLss] - S[17]) * exp(-dt / C[K_thL]);
  S[18] = A[A_hLssp] - (A[A_hLssp] - S[18]) * exp(-dt / C[K_thLp]);
  RL(19, A_ass, A_ta);
  RL(20, A_iss, A_tiF);
  RL(21, A_iss, A_tiS);
  RL(22, A_assp, A_ta);
  RL(23, A_iss, A_tiFp);
  RL(24, A_iss, A_tiSp);
  RL(25, A_dss, A_td);
  RL(26, A_fss, A_tff);
  RL(27, A_fss, A_tfs);
  RL(28, A_fcass, A_tfcaf);
  RL(29, A_fcass, A_tfcas);
  S[30] = A[A_fcass] - (A[A_fcass] - S[30]) * exp(-dt / C[K_tjca]);
  RL(31, A_fss, A_tffp);
  /* fcafp: the reference divides the STATES index of d (25) by tfcafp, not dt */
  S[32] = A[A_fcass] - (A[A_fcass] - S[32]) * exp(-25.0 / A[A_tfcafp]);
  {
    const double ninf = A[A_anca] * C[K_k2n] / A[A_km2n];
    S[33] = ninf - (ninf - S[33]) * exp(-A[A_km2n] * dt);
  }
  RL(34, A_xrss, A_txrf);
  RL(35, A_xrss, A_txrs);
  RL(36, A_xs1ss, A_txs1);
  RL(37, A_xs2ss, A_txs2);
  RL(38, A_xk1ss, A_txk1);
  RL(39, A_Jrel_inf, A_tau_rel);
  RL(40, A_Jrel_infp, A_tau_relp);
#undef RL
  S[1] += R[1] * dt;
  S[0] += R[0] * dt;
  S[3] += R[3] * dt;
  S[4] += R[4] * dt;
  S[5] += R[5] * dt;
  S[6] += R[6] * dt;
  S[9] += R[9] * dt;
  S[2] += R[2] * dt;
  S[7] += R[7] * dt;
  S[8] += R[8] * dt;
}

/* set_time_step (:1097-1137) */
double ord11_time_step(double t, double time_point, double max_dt, const double *C, const double *R)
{
  double ts = 0.005;
  if (t <= time_point || (t - floor(t / C[K_BCL]) * C[K_BCL]) <= time_point) return ts;
  const double dv = R[0];
  if (fabs(dv * ts) <= 0.2) {
    ts = fabs(0.8 / dv);
    if (ts < 0.005) ts = 0.005;
    else if (ts > max_dt) ts = max_dt;
  } else if (fabs(dv * ts) >= 0.8) {
    ts = fabs(0.2 / dv);
    while (fabs(dv * ts) >= 0.8 && 0.005 < ts && ts < max_dt) ts = ts / 10.0;
  }
  return ts;
}

#ifndef ORD11_ONLY_MODEL
/* One cell through the DrugSimulation ORd 2011 driver loop (drug_sim.cpp:306-420): adaptive
   Rush-Larsen, steps clipped to beat ends, and per-beat features over the last nwin beats.
   Output record res[ORD11_NBM] (see oracle.h).  Returns the number of steps taken. */
long oracle_ord11_run(const ord11_cell_cfg_t *cfg, double *states_io, int use_states_in, double *res)
{
  double C[ORD11_NC], S[ORD11_NS], R[ORD11_NS], A[ORD11_NA];
  memset(A, 0, sizeof(A));
  memset(R, 0, sizeof(R));
  ord11_init_consts(C, S, cfg->celltype);
  if (cfg->dutta) ord11_apply_dutta(C);
  ord11_apply_drug(C, cfg->conc, cfg->ic50);
  C[K_BCL] = cfg->bcl;
  if (use_states_in) memcpy(S, states_io, sizeof(S));
  const double bcl = cfg->bcl, tmax = cfg->nbeats * bcl;
  const int first_win = cfg->nbeats - cfg->nwin;
  double tcurr = 0.0, dt = 0.005;
  int pace = 0, steep = 0, eligible = 0;
  double r30 = 0, r50 = 0, r90 = 0, t_peak = 0;
  double qap = 0, q4ap = 0, inal_ap = 0, ical_ap = 0, qcl = 0, q4cl = 0, inal_cl = 0, ical_cl = 0;
  /* per-beat record and the selected one: dvmdt_repol, vm_peak, vm_valley + the integrals */
  double tmp[ORD11_NBM], best[ORD11_NBM];
  memset(tmp, 0, sizeof(tmp));
  tmp[ORD11_BM_DVMDT_REPOL] = -999; tmp[ORD11_BM_VM_PEAK] = -999; tmp[ORD11_BM_VM_VALLEY] = S[0];
  memcpy(best, tmp, sizeof(tmp));
  long nsteps = 0;
  while (tcurr < tmax) {
    ord11_rates(tcurr, C, R, S, A);
    const double dts = ord11_time_step(tcurr, cfg->time_point, cfg->max_dt, C, R);
    if (floor((tcurr + dts) / bcl) == floor(tcurr / bcl)) {
      dt = dts;
    } else {
      dt = (floor(tcurr / bcl) + 1) * bcl - tcurr;
      if (eligible && pace >= first_win) {
        tmp[ORD11_BM_QNET_AP] = qap; tmp[ORD11_BM_QNET4_AP] = q4ap;
        tmp[ORD11_BM_INAL_AUC_AP] = inal_ap; tmp[ORD11_BM_ICAL_AUC_AP] = ical_ap;
        tmp[ORD11_BM_QNET_CL] = qcl; tmp[ORD11_BM_QNET4_CL] = q4cl;
        tmp[ORD11_BM_INAL_AUC_CL] = inal_cl; tmp[ORD11_BM_ICAL_AUC_CL] = ical_cl;
        tmp[ORD11_BM_T_PEAK] = t_peak; tmp[ORD11_BM_VM_REPOL30] = r30;
        tmp[ORD11_BM_VM_REPOL50] = r50; tmp[ORD11_BM_VM_REPOL90] = r90;
        if (tmp[ORD11_BM_DVMDT_REPOL] > best[ORD11_BM_DVMDT_REPOL]) {
          steep = pace;
          memcpy(best, tmp, sizeof(tmp));
        }
      }
      qap = q4ap = inal_ap = ical_ap = qcl = q4cl = inal_cl = ical_cl = 0.0;
      t_peak = 0.0;
      memset(tmp, 0, sizeof(tmp));
      tmp[ORD11_BM_DVMDT_REPOL] = -999; tmp[ORD11_BM_VM_PEAK] = -999; tmp[ORD11_BM_VM_VALLEY] = S[0];
      pace++;
      eligible = 0;
    }
    ord11_rush_larsen(dt, C, R, S, A);
    ++nsteps;
    if (pace >= first_win) {
      const double t_beat = bcl * pace + C[K_stim_start];
      if (tcurr > t_beat + 2 && tcurr < t_beat + 10 && fabs(A[A_INa]) < 1) {
        if (S[0] > tmp[ORD11_BM_VM_PEAK]) {
          tmp[ORD11_BM_VM_PEAK] = S[0];
          if (tmp[ORD11_BM_VM_PEAK] > 0) {
            const double amp = tmp[ORD11_BM_VM_PEAK] - tmp[ORD11_BM_VM_VALLEY];
            r30 = tmp[ORD11_BM_VM_PEAK] - 0.3 * amp;
            r50 = tmp[ORD11_BM_VM_PEAK] - 0.5 * amp;
            r90 = tmp[ORD11_BM_VM_PEAK] - 0.9 * amp;
            eligible = 1;
            t_peak = tcurr;
          } else {
            eligible = 0;
          }
        }
      } else if (tcurr > t_beat + 10 && eligible) {
        if (R[0] > tmp[ORD11_BM_DVMDT_REPOL] && S[0] <= r30 && S[0] >= r90) tmp[ORD11_BM_DVMDT_REPOL] = R[0];
      }
      const double inet = A[A_INaL] + A[A_ICaL] + A[A_Ito] + A[A_IKr] + A[A_IKs] + A[A_IK1];
      const double inet4 = A[A_INaL] + A[A_ICaL] + A[A_IKr] + A[A_INa];
      if (eligible && S[0] > r90) {
        qap += (inet * dt) / 1000.;
        q4ap += (inet4 * dt) / 1000.;
        inal_ap += A[A_INaL] * dt;
        ical_ap += A[A_ICaL] * dt;
      }
      qcl += (inet * dt) / 1000.;
      q4cl += (inet4 * dt) / 1000.;
      inal_cl += A[A_INaL] * dt;
      ical_cl += A[A_ICaL] * dt;
    }
    tcurr += dt;
  }
  memcpy(res, best, sizeof(best));
  res[ORD11_BM_STEEPEST_PACE] = steep;
  if (states_io) memcpy(states_io, S, sizeof(S));
  return nsteps;
}

/* ncell independent cells over nthreads host threads (round-robin, as the reference spreads
   samples over MPI ranks); states_in: ncell x 41 start states or NULL; res: ncell x 16;
   states_out: ncell x 41 or NULL.  Returns the wall seconds. */
typedef struct {
  const ord11_cell_cfg_t *cfgs;
  const double *sin;
  double *res, *sout;
  int ncell, tid, nthreads;
} ord11_job_t;

static void *ord11_worker(void *p)
{
  ord11_job_t *j = (ord11_job_t *)p;
  for (int i = j->tid; i < j->ncell; i += j->nthreads) {
    double S[ORD11_NS];
    if (j->sin) memcpy(S, j->sin + (size_t)i * ORD11_NS, sizeof(S));
    oracle_ord11_run(&j->cfgs[i], S, j->sin != NULL, j->res + (size_t)i * ORD11_NBM);
    if (j->sout) memcpy(j->sout + (size_t)i * ORD11_NS, S, sizeof(S));
  }
  return NULL;
}

double oracle_ord11_run_many(const ord11_cell_cfg_t *cfgs, int ncell, int nthreads, const double *states_in,
                             double *res, double *states_out)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  ord11_job_t jobs[64];
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int t = 0; t < nthreads; ++t) {
    jobs[t] = (ord11_job_t){ cfgs, states_in, res, states_out, ncell, t, nthreads };
    pthread_create(&th[t], NULL, ord11_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &b);
  return (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
}
#endif
