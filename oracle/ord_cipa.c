/*
 * ord_cipa.c -- ORACLE (test infrastructure only): FP64 restatement of the
 * O'Hara-Rudy CiPA v1 2017 ventricular model with the Markov IKr/drug-binding
 * chain, as exported by CellML in
 *   CEP5Cell/cellmodels/ohara_rudy_cipa_v1_2017.cpp
 * (identical to CEP7Cell; state names from CEP7Cell/enums/enum_ohara_rudy_cipa_v1_2017.hpp).
 *
 * Array layout is the reference's: STATES[52], CONSTANTS[206], RATES[52] and the
 * ALGEBRAIC slots that the reference drivers read (currents).  Formulas are written
 * with named locals, in the reference's evaluation order so that FP64 results agree
 * to rounding.  Pinned by tests/golden/ord_*.bin (compiled reference).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

/* initial conditions + parameters: ___initConsts, ohara_rudy_cipa_v1_2017.cpp:545-804 */
void ord_init_consts(double *C, double *S, double celltype)
{
  memset(C, 0, sizeof(double) * ORD_NC);
  memset(S, 0, sizeof(double) * ORD_NS);
  C[0] = celltype;
  /* extracellular ions, physical constants, geometry (:547-557) */
  C[1] = 140; C[2] = 1.8; C[3] = 5.4; C[4] = 8314; C[5] = 310; C[6] = 96485;
  C[7] = 1; C[8] = 2; C[9] = 1; C[10] = 0.01; C[11] = 0.0011;
  /* stimulus (:559-563): start, end, amplitude, period, duration */
  C[12] = 3; C[13] = 100000000000000000.0; C[14] = -80; C[15] = 1000; C[16] = 0.5;
  /* CaMK (:564-568) */
  C[17] = 0.15; C[18] = 0.05; C[19] = 0.00068; C[20] = 0.05; C[21] = 0.0015;
  /* buffers (:571-580) */
  C[22] = 0.05; C[23] = 0.00238; C[24] = 0.07; C[25] = 0.0005; C[26] = 0.047;
  C[27] = 0.00087; C[28] = 1.124; C[29] = 0.0087; C[30] = 10; C[31] = 0.8;
  C[32] = 1; C[33] = 0.01833;
  /* INa (:590-607) */
  C[34] = 39.57; C[35] = 9.871; C[36] = 11.64; C[37] = 34.77; C[38] = 6.765;
  C[39] = 8.552; C[40] = 77.42; C[41] = 5.955; C[42] = 82.9; C[43] = 6.086;
  C[44] = 0.99; C[45] = 75; C[46] = 0;
  /* INaL, Ito, ICaL, IKr (:613-631) */
  C[47] = 200; C[48] = 0.019957499999999975; C[49] = 0.02;
  C[50] = 0.002; C[51] = 1000; C[52] = 0.0001007; C[53] = 0.04658545454545456;
  /* Markov IKr rate parameters A,B,q triplets (:651-692) */
  {
    static const double mk[42] = {
      0.0264, 4.631E-05, 4.843,      4.986E-06, -0.004226, 4.23,
      0.001214, 0.008516, 4.962,     1.854E-05, -0.04641, 3.769,
      0.0007868, 1.535E-08, 4.942,   5.455E-06, -0.1688, 4.156,
      0.005509, 7.771E-09, 4.22,     0.001416, -0.02877, 1.459,
      0.4492, 0.008595, 5,           0.3181, 3.613E-08, 4.663,
      0.149, 0.004668, 2.412,        0.01241, 0.1725, 5.568,
      0.3226, -0.0006575, 5,         0.008978, -0.02215, 5.682 };
    memcpy(C + 54, mk, sizeof(mk));
  }
  /* dyn-hERG drug parameters Kmax, Ku, n, halfmax, Kt, Vhalf, Temp (:693-699) */
  C[96] = 0; C[97] = 0; C[98] = 1; C[99] = 1; C[100] = 0; C[101] = 1; C[102] = 37;
  /* IKs, IK1 (:700-706) */
  C[103] = 0.006358000000000001; C[104] = 817.3; C[105] = 0.3239783999999998;
  /* INaCa (:707-716) */
  C[106] = 15; C[107] = 5; C[108] = 88.12; C[109] = 12.5; C[110] = 6e4; C[111] = 6e4;
  C[112] = 5e3; C[113] = 1.5e6; C[114] = 5e3; C[115] = 0.5224; C[116] = 0.167;
  C[117] = 150e-6; C[118] = 0.0008;
  /* INaK (:719-736) */
  C[119] = 949.5; C[120] = 182.4; C[121] = 687.2; C[122] = 39.4; C[123] = 1899;
  C[124] = 79300; C[125] = 639; C[126] = 40; C[127] = 9.073; C[128] = 27.78;
  C[129] = -0.155; C[130] = 0.5; C[131] = 0.3582; C[132] = 0.05; C[133] = 9.8;
  C[134] = 1.698e-7; C[135] = 1e-7; C[136] = 4.2; C[137] = 1.698e-7; C[138] = 224;
  C[139] = 292;
  /* IKb, INab, ICab, IpCa, Jrel (:737-743) */
  C[140] = 30; C[141] = 0.003; C[142] = 3.75e-10; C[143] = 2.5e-8; C[144] = 0.0005;
  C[145] = 0.0005; C[146] = 4.75;
  C[147] = 1.0; C[148] = 1.0;

  /* initial states (:558-748) */
  {
    static const double s0[49] = {
      -88.00190465, 0.0125840447, 8.49e-05, 7.268004498, 7.268089977, 144.6555918,
      144.6555651, 1.619574538, 1.571234014, 8.6e-05,
      0.007344121102, 0.6981071913, 0.6980895801, 0.6979908432, 0.4549485525,
      0.6979245865, 0.0001882617273, 0.5008548855, 0.2693065357,
      0.001001097687, 0.9995541745, 0.5865061736, 0.0005100862934, 0.9995541823,
      0.6393399482,
      2.34e-9, 0.9999999909, 0.9102412777, 0.9999999909, 0.9998046777, 0.9999738312,
      0.9999999909, 0.9999999909, 0.002749414044,
      0.999637, 6.83208e-05, 1.80145e-08, 8.26619e-05, 0.00015551, 5.67623e-05,
      0, 0, 0, 0,
      0.2707758025, 0.0001928503426, 0.9967597594, 2.5e-7, 3.12e-7 };
    memcpy(S, s0, sizeof(s0));
    S[49] = S[50] = S[51] = 0.;
  }

  /* derived constants (:749-803), celltype 0 endo / 1 epi / 2 M */
  const int epi = (C[0] == 1.0), mid = (C[0] == 2.0);
  C[149] = C[6] / (C[4] * C[5]);
  C[150] = epi ? C[22] * 1.3 : C[22];
  C[151] = 1.0 - C[44];
  C[152] = 3.0 * C[47];
  C[153] = epi ? C[48] * 0.6 : C[48];
  C[154] = (epi || mid) ? C[49] * 4.0 : C[49];
  C[155] = 0.6;
  C[156] = epi ? C[52] * 1.2 : mid ? C[52] * 2.5 : C[52];
  C[157] = 75.0;
  C[158] = 0.0;
  C[159] = epi ? C[53] * 1.3 : mid ? C[53] * 0.8 : C[53];
  C[160] = epi ? C[103] * 1.4 : C[103];
  C[161] = epi ? C[105] * 1.2 : mid ? C[105] * 1.3 : C[105];
  C[162] = 1000.0 * 3.14 * C[11] * C[11] * C[10];
  C[163] = epi ? C[141] * 0.6 : C[141];
  C[164] = 0.0; C[165] = 0.0;
  C[166] = 0.5 * C[146];
  C[167] = 1.25 * C[146];
  C[168] = epi ? 1.3 : 1.0;
  C[205] = 0.0;
  C[169] = C[6] * C[149];
  C[170] = 1.0 - C[155];
  C[171] = 1.1 * C[156];
  C[172] = 0.00125 * C[156];
  C[173] = 0.0003574 * C[156];
  C[174] = 2.0 * C[149];
  C[175] = C[149]; C[176] = C[149];
  C[177] = 2.0 * 3.14 * C[11] * C[11] + 2.0 * 3.14 * C[11] * C[10];
  C[178] = C[149];
  C[179] = 2.0 * C[149];
  C[180] = 0.5 * C[167];
  C[181] = 0.00125 * C[171];
  C[182] = 0.0003574 * C[171];
  C[183] = 2.0 * C[177];
  C[184] = 0.68 * C[162];
  C[185] = 0.0552 * C[162];
  C[186] = 0.0048 * C[162];
  C[187] = 0.02 * C[162];
  C[188] = C[109] + 1.0 + (C[1] / C[106]) * (1.0 + C[1] / C[107]);
  C[189] = (C[1] * C[1]) / (C[188] * C[106] * C[107]);
  C[190] = 1.0 / C[188];
  C[191] = C[190] * C[2] * C[113];
  C[192] = C[114]; C[193] = C[114];
  C[194] = epi ? C[118] * 1.1 : mid ? C[118] * 1.4 : C[118];
  C[195] = C[188]; C[196] = C[189]; C[197] = C[190]; C[198] = C[191];
  C[199] = C[114]; C[200] = C[114];
  C[201] = C[120] * C[132];
  C[202] = C[121];
  C[203] = ((C[125] * C[133]) / C[134]) / (1.0 + C[133] / C[134]);
  C[204] = epi ? C[140] * 0.9 : mid ? C[140] * 0.7 : C[140];
}

/* Hill block factor as written in initConsts (:826-837): "NA" / zero IC50 -> no block */
static double hill_factor(double D, double ic50, double h)
{
  return (ic50 > 10E-14 && h > 10E-14) ? pow(1. + pow(D / ic50, h), -1) : 1.;
}

/* initConsts(celltype, conc, herg, drug), ohara_rudy_cipa_v1_2017.cpp:819-845.
   Note the reference quirk kept here: PCap/PCaNa/PCaK (C[171..173,181,182]) were derived
   before this scaling, so only the non-phosphorylated ICaL permeability is blocked. */
void ord_apply_drug(double *C, double *S, double conc, const double *herg)
{
  S[43] = conc;
  if (!herg) return;
  C[156] *= hill_factor(S[43], herg[0], herg[1]);
  C[161] *= hill_factor(S[43], herg[2], herg[3]);
  C[160] *= hill_factor(S[43], herg[4], herg[5]);
  C[45]  *= hill_factor(S[43], herg[6], herg[7]);
  C[153] *= hill_factor(S[43], herg[8], herg[9]);
  C[154] *= hill_factor(S[43], herg[10], herg[11]);
  C[96] = herg[14]; C[97] = herg[15]; C[98] = herg[16]; C[99] = herg[17]; C[101] = herg[18];
}

/* GHK-type flux with the reference's |U| <= 1e-7 guard */
static double ghk(double Apre, double U)
{
  return (-1.0e-07 <= U && U <= 1.0e-07) ? Apre * (1.0 - 0.5 * U) : (Apre * U) / (exp(U) - 1.0);
}

/* Markov rate constant A*exp(B*V)*Q10-factor, factor as exp((Temp-20)*log(q)/10) */
#define MK(i, v) (C[(i)] * exp(C[(i) + 1] * (v)))
#define Q10(i) exp(((C[102] - 20.0) * log(C[(i) + 2])) / 10.0)

/* computeRates, ohara_rudy_cipa_v1_2017.cpp:847-1111 (SINGLE_CELL stimulus) */
void ord_rates(double t, const double *C, double *R, const double *S, double *A)
{
  const double v = S[0], camkt = S[1], cass = S[2], nai = S[3], nass = S[4], ki = S[5],
               kss = S[6], cansr = S[7], cajsr = S[8], cai = S[9];
  const double m = S[10], hf = S[11], hs = S[12], j = S[13], hsp = S[14], jp = S[15],
               mL = S[16], hL = S[17], hLp = S[18], a = S[19], iF = S[20], iS = S[21],
               ap = S[22], iFp = S[23], iSp = S[24], d = S[25], ff = S[26], fs = S[27],
               fcaf = S[28], fcas = S[29], jca = S[30], ffp = S[31], fcafp = S[32], nca = S[33];
  const double IC1 = S[34], IC2 = S[35], C1 = S[36], C2 = S[37], O = S[38], IO = S[39],
               IObound = S[40], Obound = S[41], Cbound = S[42], D = S[43];
  const double xs1 = S[44], xs2 = S[45], xk1 = S[46], Jrelnp = S[47], Jrelp = S[48];
  const double celltype = C[0], nao = C[1], cao = C[2], ko = C[3], Rg = C[4], T = C[5], F = C[6];
  const double shift = C[46];

  /* gate steady states and time constants (:849-887) */
  const double hLss = 1.0 / (1.0 + exp((v + 87.61) / 7.488));
  const double hLssp = 1.0 / (1.0 + exp((v + 93.81) / 7.488));
  const double mss = 1.0 / (1.0 + exp(-(v + C[34]) / C[35]));
  const double tm = 1.0 / (C[38] * exp((v + C[36]) / C[37]) + C[39] * exp(-(v + C[40]) / C[41]));
  const double hss = 1.0 / (1.0 + exp(((v + C[42]) - shift) / C[43]));
  const double thf = 1.0 / (1.432e-05 * exp(-((v + 1.196) - shift) / 6.285) + 6.149 * exp(((v + 0.5096) - shift) / 20.27));
  const double ths = 1.0 / (0.009794 * exp(-((v + 17.95) - shift) / 28.05) + 0.3343 * exp(((v + 5.73) - shift) / 56.66));
  const double ass = 1.0 / (1.0 + exp(-(v - 14.34) / 14.82));
  const double ta = 1.0515 / (1.0 / (1.2089 * (1.0 + exp(-(v - 18.4099) / 29.3814))) + 3.5 / (1.0 + exp((v + 100.0) / 29.3814)));
  const double dss = 1.0 / (1.0 + exp(-(v + 3.94) / 4.23));
  const double td = 0.6 + 1.0 / (exp(-0.05 * (v + 6.0)) + exp(0.09 * (v + 14.0)));
  const double fss = 1.0 / (1.0 + exp((v + 19.58) / 3.696));
  const double tff = 7.0 + 1.0 / (0.0045 * exp(-(v + 20.0) / 10.0) + 0.0045 * exp((v + 20.0) / 10.0));
  const double tfs = 1000.0 + 1.0 / (3.5e-05 * exp(-(v + 5.0) / 4.0) + 3.5e-05 * exp((v + 5.0) / 6.0));
  const double fcass = fss;
  const double km2n = jca * 1.0;
  const double anca = 1.0 / (C[51] / km2n + pow(1.0 + C[50] / cass, 4.0));
  const double xs1ss = 1.0 / (1.0 + exp(-(v + 11.6) / 8.932));
  const double txs1 = C[104] + 1.0 / (0.0002326 * exp((v + 48.28) / 17.8) + 0.001292 * exp(-(v + 210.0) / 230.0));
  const double xk1ss = 1.0 / (1.0 + exp(-(v + 2.5538 * ko + 144.59) / (1.5692 * ko + 3.8115)));
  const double txk1 = 122.2 / (exp(-(v + 127.2) / 20.36) + exp((v + 236.8) / 69.33));
  const double CaMKb = (C[20] * (1.0 - camkt)) / (1.0 + C[21] / cass);
  const double jss = hss;
  const double tj = 2.038 + 1.0 / (0.02136 * exp(-((v + 100.6) - shift) / 8.281) + 0.3052 * exp(((v + 0.9941) - shift) / 38.45));
  const double assp = 1.0 / (1.0 + exp(-(v - 24.34) / 14.82));
  const double tfcaf = 7.0 + 1.0 / (0.04 * exp(-(v - 4.0) / 7.0) + 0.04 * exp((v - 4.0) / 7.0));
  const double tfcas = 100.0 + 1.0 / (0.00012 * exp(-v / 3.0) + 0.00012 * exp(v / 7.0));
  const double tffp = 2.5 * tff;
  const double xs2ss = xs1ss;
  const double txs2 = 1.0 / (0.01 * exp((v - 50.0) / 20.0) + 0.0193 * exp(-(v + 66.54) / 31.0));
  const double hssp = 1.0 / (1.0 + exp(((v + 89.1) - shift) / 6.086));
  const double thsp = 3.0 * ths;
  const double tjp = 1.46 * tj;
  const double mLss = 1.0 / (1.0 + exp(-(v + 42.85) / 5.264));
  const double tmL = tm;
  const double tfcafp = 2.5 * tfcaf;
  const double iss = 1.0 / (1.0 + exp((v + 43.94) / 5.711));
  const double delta_epi = (celltype == 1.0) ? 1.0 - 0.95 / (1.0 + exp((v + 70.0) / 5.0)) : 1.0;
  const double tiF_b = 4.562 + 1.0 / (0.3933 * exp(-(v + 100.0) / 100.0) + 0.08004 * exp((v + 50.0) / 16.59));
  const double tiF = tiF_b * delta_epi;
  const double tiS_b = 23.62 + 1.0 / (0.001416 * exp(-(v + 96.52) / 59.05) + 1.78e-08 * exp((v + 114.1) / 8.079));
  const double tiS = tiS_b * delta_epi;
  const double dti_develop = 1.354 + 0.0001 / (exp((v - 167.4) / 15.89) + exp(-(v - 12.23) / 0.2154));
  const double dti_recover = 1.0 - 0.5 / (1.0 + exp((v + 70.0) / 20.0));
  const double tiFp = dti_develop * dti_recover * tiF;
  const double tiSp = dti_develop * dti_recover * tiS;

  /* ICaL and its Na/K components (:888-896, :962-965, :1003-1006) */
  const double f = C[155] * ff + C[170] * fs;
  const double Afcaf = 0.3 + 0.6 / (1.0 + exp((v - 10.0) / 10.0));
  const double Afcas = 1.0 - Afcaf;
  const double fca = Afcaf * fcaf + Afcas * fcas;
  const double fp = C[155] * ffp + C[170] * fs;
  const double fcap = Afcaf * fcafp + Afcas * fcas;
  const double vfrt = v * C[149];
  const double PhiCaL = ghk((4.0 * C[169] * (cass * exp(2.0 * vfrt) - 0.341 * cao)) / C[174], C[174] * (v - C[158]));
  const double CaMKa = CaMKb + camkt;
  const double fICaLp = 1.0 / (1.0 + C[17] / CaMKa);
  const double ICaL = (1.0 - fICaLp) * C[156] * PhiCaL * d * (f * (1.0 - nca) + jca * fca * nca)
                    + fICaLp * C[171] * PhiCaL * d * (fp * (1.0 - nca) + jca * fcap * nca);

  /* Jrel targets (:897-904) */
  const double Jrel_inf_b = (C[166] * -ICaL) / (1.0 + 1.0 * pow(1.5 / cajsr, 8.0));
  const double Jrel_inf = (celltype == 2.0) ? Jrel_inf_b * 1.7 : Jrel_inf_b;
  const double tau_rel_b = C[146] / (1.0 + 0.0123 / cajsr);
  const double tau_rel = (tau_rel_b < 0.001) ? 0.001 : tau_rel_b;
  const double Jrel_infp_b = (C[180] * -ICaL) / (1.0 + pow(1.5 / cajsr, 8.0));
  const double Jrel_infp = (celltype == 2.0) ? Jrel_infp_b * 1.7 : Jrel_infp_b;
  const double tau_relp_b = C[167] / (1.0 + 0.0123 / cajsr);
  const double tau_relp = (tau_relp_b < 0.001) ? 0.001 : tau_relp_b;

  /* K currents (:905-918) */
  const double EK = ((Rg * T) / F) * log(ko / ki);
  const double AiF = 1.0 / (1.0 + exp((v - 213.6) / 151.2));
  const double AiS = 1.0 - AiF;
  const double ii = AiF * iF + AiS * iS;
  const double ip = AiF * iFp + AiS * iSp;
  const double fItop = 1.0 / (1.0 + C[17] / CaMKa);
  const double Ito = C[154] * (v - EK) * ((1.0 - fItop) * a * ii + fItop * ap * ip);
  const double IKr = C[159] * pow(ko / 5.4, 1.0 / 2) * O * (v - EK);
  const double EKs = ((Rg * T) / F) * log((ko + C[33] * nao) / (ki + C[33] * nai));
  const double KsCa = 1.0 + 0.6 / (1.0 + pow(3.8e-05 / cai, 1.4));
  const double IKs = C[160] * KsCa * xs1 * xs2 * (v - EKs);
  const double rk1 = 1.0 / (1.0 + exp(((v + 105.8) - 2.6 * ko) / 9.493));
  const double IK1 = C[161] * pow(ko, 1.0 / 2) * rk1 * xk1 * (v - EK);

  /* INaK (:919-941) */
  const double Knao = C[128] * exp(((1.0 - C[129]) * v * F) / (3.0 * Rg * T));
  const double a3 = (C[123] * pow(ko / C[131], 2.0)) / ((pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / C[131], 2.0)) - 1.0);
  const double Pnak_P = C[136] / (1.0 + C[135] / C[137] + nai / C[138] + ki / C[139]);
  const double b3 = (C[124] * Pnak_P * C[135]) / (1.0 + C[133] / C[134]);
  const double Knai = C[127] * exp((C[129] * v * F) / (3.0 * Rg * T));
  const double a1 = (C[119] * pow(nai / Knai, 3.0)) / ((pow(1.0 + nai / Knai, 3.0) + pow(1.0 + ki / C[130], 2.0)) - 1.0);
  const double b2 = (C[122] * pow(nao / Knao, 3.0)) / ((pow(1.0 + nao / Knao, 3.0) + pow(1.0 + ko / C[131], 2.0)) - 1.0);
  const double b4 = (C[126] * pow(ki / C[130], 2.0)) / ((pow(1.0 + nai / Knai, 3.0) + pow(1.0 + ki / C[130], 2.0)) - 1.0);
  const double a2 = C[202], a4 = C[203], b1 = C[201];
  const double xk1_ = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
  const double xk2_ = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
  const double xk3_ = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
  const double xk4_ = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
  const double Ek1 = xk1_ / (xk1_ + xk2_ + xk3_ + xk4_);
  const double Ek2 = xk2_ / (xk1_ + xk2_ + xk3_ + xk4_);
  const double JnakNa = 3.0 * (Ek1 * a3 - Ek2 * b3);
  const double Ek3 = xk3_ / (xk1_ + xk2_ + xk3_ + xk4_);
  const double Ek4 = xk4_ / (xk1_ + xk2_ + xk3_ + xk4_);
  const double JnakK = 2.0 * (Ek4 * b1 - Ek3 * a1);
  const double INaK = C[204] * (C[7] * JnakNa + C[9] * JnakK);

  const double xkb = 1.0 / (1.0 + exp(-(v - 14.48) / 18.34));
  const double IKb = C[163] * xkb * (v - EK);
  const double Istim = (t >= C[12] && t <= C[13] &&
                        (t - C[12]) - floor((t - C[12]) / C[15]) * C[15] <= C[16]) ? C[14] : 0.0;
  const double JdiffK = (kss - ki) / 2.0;
  const double PhiCaK = ghk((0.75 * C[169] * (kss * exp(vfrt) - ko)) / C[176], C[176] * (v - C[158]));
  const double ICaK = (1.0 - fICaLp) * C[173] * PhiCaK * d * (f * (1.0 - nca) + jca * fca * nca)
                    + fICaLp * C[182] * PhiCaK * d * (fp * (1.0 - nca) + jca * fcap * nca);

  /* Na currents (:966-973) */
  const double ENa = ((Rg * T) / F) * log(nao / nai);
  const double hh = C[44] * hf + C[151] * hs;
  const double hp = C[44] * hf + C[151] * hsp;
  const double fINap = 1.0 / (1.0 + C[17] / CaMKa);
  const double INa = C[45] * (v - ENa) * pow(m, 3.0) * ((1.0 - fINap) * hh * j + fINap * hp * jp);
  const double fINaLp = 1.0 / (1.0 + C[17] / CaMKa);
  const double INaL = C[153] * (v - ENa) * mL * ((1.0 - fINaLp) * hL + fINaLp * hLp);

  /* INaCa_i (:974-1001) */
  const double allo_i = 1.0 / (1.0 + pow(C[117] / cai, 2.0));
  const double hna = exp((C[115] * v * F) / (Rg * T));
  const double h7_i = 1.0 + (nao / C[108]) * (1.0 + 1.0 / hna);
  const double h8_i = nao / (C[108] * hna * h7_i);
  const double k3pp_i = h8_i * C[112];
  const double h1_i = 1.0 + (nai / C[108]) * (1.0 + hna);
  const double h2_i = (nai * hna) / (C[108] * h1_i);
  const double k4pp_i = h2_i * C[112];
  const double h4_i = 1.0 + (nai / C[106]) * (1.0 + nai / C[107]);
  const double h5_i = (nai * nai) / (h4_i * C[106] * C[107]);
  const double k7_i = h5_i * h2_i * C[110];
  const double k8_i = h8_i * C[189] * C[110];
  const double h9_i = 1.0 / h7_i;
  const double k3p_i = h9_i * C[111];
  const double k3_i = k3p_i + k3pp_i;
  const double hca = exp((C[116] * v * F) / (Rg * T));
  const double h3_i = 1.0 / h1_i;
  const double k4p_i = (h3_i * C[111]) / hca;
  const double k4_i = k4p_i + k4pp_i;
  const double h6_i = 1.0 / h4_i;
  const double k6_i = h6_i * cai * C[113];
  const double x1_i = C[192] * k4_i * (k7_i + k6_i) + C[193] * k7_i * (C[192] + k3_i);
  const double x2_i = C[191] * k7_i * (k4_i + C[193]) + k4_i * k6_i * (C[191] + k8_i);
  const double x3_i = C[191] * k3_i * (k7_i + k6_i) + k8_i * k6_i * (C[192] + k3_i);
  const double x4_i = C[192] * k8_i * (k4_i + C[193]) + k3_i * C[193] * (C[191] + k8_i);
  const double E1_i = x1_i / (x1_i + x2_i + x3_i + x4_i);
  const double E2_i = x2_i / (x1_i + x2_i + x3_i + x4_i);
  const double E3_i = x3_i / (x1_i + x2_i + x3_i + x4_i);
  const double E4_i = x4_i / (x1_i + x2_i + x3_i + x4_i);
  const double JncxNa_i = (3.0 * (E4_i * k7_i - E1_i * k8_i) + E3_i * k4pp_i) - E2_i * k3pp_i;
  const double JncxCa_i = E2_i * C[192] - E1_i * C[191];
  const double INaCa_i = 0.8 * C[194] * allo_i * (C[7] * JncxNa_i + C[8] * JncxCa_i);

  const double PhiNab = ghk((C[142] * C[169] * (nai * exp(vfrt) - nao)) / C[178], C[178] * (v - C[164]));
  const double INab = PhiNab;
  const double JdiffNa = (nass - nai) / 2.0;
  const double PhiCaNa = ghk((0.75 * C[169] * (nass * exp(vfrt) - nao)) / C[175], C[175] * (v - C[158]));
  const double ICaNa = (1.0 - fICaLp) * C[172] * PhiCaNa * d * (f * (1.0 - nca) + jca * fca * nca)
                     + fICaLp * C[181] * PhiCaNa * d * (fp * (1.0 - nca) + jca * fcap * nca);

  /* INaCa_ss (:1007-1034) */
  const double allo_ss = 1.0 / (1.0 + pow(C[117] / cass, 2.0));
  const double h7_s = 1.0 + (nao / C[108]) * (1.0 + 1.0 / hna);
  const double h8_s = nao / (C[108] * hna * h7_s);
  const double k3pp_s = h8_s * C[112];
  const double h1_s = 1.0 + (nass / C[108]) * (1.0 + hna);
  const double h2_s = (nass * hna) / (C[108] * h1_s);
  const double k4pp_s = h2_s * C[112];
  const double h4_s = 1.0 + (nass / C[106]) * (1.0 + nass / C[107]);
  const double h5_s = (nass * nass) / (h4_s * C[106] * C[107]);
  const double k7_s = h5_s * h2_s * C[110];
  const double k8_s = h8_s * C[196] * C[110];
  const double h9_s = 1.0 / h7_s;
  const double k3p_s = h9_s * C[111];
  const double k3_s = k3p_s + k3pp_s;
  const double h3_s = 1.0 / h1_s;
  const double k4p_s = (h3_s * C[111]) / hca;
  const double k4_s = k4p_s + k4pp_s;
  const double h6_s = 1.0 / h4_s;
  const double k6_s = h6_s * cass * C[113];
  const double x1_s = C[199] * k4_s * (k7_s + k6_s) + C[200] * k7_s * (C[199] + k3_s);
  const double x2_s = C[198] * k7_s * (k4_s + C[200]) + k4_s * k6_s * (C[198] + k8_s);
  const double x3_s = C[198] * k3_s * (k7_s + k6_s) + k8_s * k6_s * (C[199] + k3_s);
  const double x4_s = C[199] * k8_s * (k4_s + C[200]) + k3_s * C[200] * (C[198] + k8_s);
  const double E1_s = x1_s / (x1_s + x2_s + x3_s + x4_s);
  const double E2_s = x2_s / (x1_s + x2_s + x3_s + x4_s);
  const double E3_s = x3_s / (x1_s + x2_s + x3_s + x4_s);
  const double E4_s = x4_s / (x1_s + x2_s + x3_s + x4_s);
  const double JncxNa_s = (3.0 * (E4_s * k7_s - E1_s * k8_s) + E3_s * k4pp_s) - E2_s * k3pp_s;
  const double JncxCa_s = E2_s * C[199] - E1_s * C[198];
  const double INaCa_ss = 0.2 * C[194] * allo_ss * (C[7] * JncxNa_s + C[8] * JncxCa_s);

  /* Ca pump/background, fluxes and buffers (:1035-1053) */
  const double IpCa = (C[144] * cai) / (C[145] + cai);
  const double ICab = ghk((C[143] * 4.0 * C[169] * (cai * exp(2.0 * vfrt) - 0.341 * cao)) / C[179], C[179] * (v - C[165]));
  const double Jdiff = (cass - cai) / 0.2;
  const double fJrelp = 1.0 / (1.0 + C[17] / CaMKa);
  const double Jrel = C[147] * ((1.0 - fJrelp) * Jrelnp + fJrelp * Jrelp);
  const double Bcass = 1.0 / (1.0 + (C[26] * C[27]) / pow(C[27] + cass, 2.0) + (C[28] * C[29]) / pow(C[29] + cass, 2.0));
  const double Jupnp = (C[168] * 0.004375 * cai) / (cai + 0.00092);
  const double Jupp = (C[168] * 2.75 * 0.004375 * cai) / ((cai + 0.00092) - 0.00017);
  const double fJupp = 1.0 / (1.0 + C[17] / CaMKa);
  const double Jleak = (0.0039375 * cansr) / 15.0;
  const double Jup = C[148] * (((1.0 - fJupp) * Jupnp + fJupp * Jupp) - Jleak);
  const double Bcai = 1.0 / (1.0 + (C[150] * C[23]) / pow(C[23] + cai, 2.0) + (C[24] * C[25]) / pow(C[25] + cai, 2.0));
  const double Jtr = (cansr - cajsr) / 100.0;
  const double Bcajsr = 1.0 / (1.0 + (C[30] * C[31]) / pow(C[31] + cajsr, 2.0));

  /* Markov IKr with drug binding (:1059-1068) */
  const double k1 = MK(54, v), k2 = MK(57, v), k3 = MK(60, v), k4 = MK(63, v);
  const double k11 = MK(66, v), k21 = MK(69, v), k31 = MK(72, v), k41 = MK(75, v);
  const double k51 = MK(78, v), k52 = MK(81, v), k53 = MK(84, v), k61 = MK(87, v);
  const double k62 = MK(90, v), k63 = MK(93, v);
  const double Dn = exp(C[98] * log(D));
  const double bind = (C[96] * C[97] * Dn) / (Dn + C[99]);
  const double trap = C[100] / (1.0 + exp(-(v - C[101]) / 6.789));
  const double unbindIO = (C[97] * C[84] * exp(C[85] * v) * Q10(84)) / (C[93] * exp(C[94] * v) * Q10(93));

  R[43] = C[205];
  R[34] = (-(k11 * IC1 * Q10(66) - k21 * IC2 * Q10(69)) + k51 * C1 * Q10(78)) - k61 * IC1 * Q10(87);
  R[35] = (((k11 * IC1 * Q10(66) - k21 * IC2 * Q10(69)) - (k3 * IC2 * Q10(60) - k4 * IO * Q10(63))) + k52 * C2 * Q10(81)) - k62 * IC2 * Q10(90);
  R[36] = -(k1 * C1 * Q10(54) - k2 * C2 * Q10(57)) - (k51 * C1 * Q10(78) - k61 * IC1 * Q10(87));
  R[37] = ((k1 * C1 * Q10(54) - k2 * C2 * Q10(57)) - (k31 * C2 * Q10(72) - k41 * O * Q10(75))) - (k52 * C2 * Q10(81) - k62 * IC2 * Q10(90));
  R[38] = ((k31 * C2 * Q10(72) - k41 * O * Q10(75)) - (k53 * O * Q10(84) - k63 * IO * Q10(93))) - (bind * O - C[97] * Obound);
  R[39] = (((k3 * IC2 * Q10(60) - k4 * IO * Q10(63)) + k53 * O * Q10(84)) - k63 * IO * Q10(93)) - (bind * IO - unbindIO * IObound);
  R[40] = ((bind * IO - unbindIO * IObound) + trap * Cbound) - C[100] * IObound;
  R[41] = ((bind * O - C[97] * Obound) + trap * Cbound) - C[100] * Obound;
  R[42] = -(trap * Cbound - C[100] * Obound) - (trap * Cbound - C[100] * IObound);

  /* HH gates and CaMK (:1069-1098) */
  R[17] = (hLss - hL) / C[47];
  R[18] = (hLssp - hLp) / C[152];
  R[10] = (mss - m) / tm;
  R[11] = (hss - hf) / thf;
  R[12] = (hss - hs) / ths;
  R[19] = (ass - a) / ta;
  R[25] = (dss - d) / td;
  R[26] = (fss - ff) / tff;
  R[27] = (fss - fs) / tfs;
  R[30] = (fcass - jca) / C[157];
  R[33] = anca * C[51] - nca * km2n;
  R[44] = (xs1ss - xs1) / txs1;
  R[46] = (xk1ss - xk1) / txk1;
  R[1] = C[18] * CaMKb * (CaMKb + camkt) - C[19] * camkt;
  R[13] = (jss - j) / tj;
  R[22] = (assp - ap) / ta;
  R[28] = (fcass - fcaf) / tfcaf;
  R[29] = (fcass - fcas) / tfcas;
  R[31] = (fss - ffp) / tffp;
  R[45] = (xs2ss - xs2) / txs2;
  R[14] = (hssp - hsp) / thsp;
  R[15] = (jss - jp) / tjp;
  R[16] = (mLss - mL) / tmL;
  R[32] = (fcass - fcafp) / tfcafp;
  R[20] = (iss - iF) / tiF;
  R[21] = (iss - iS) / tiS;
  R[23] = (iss - iFp) / tiFp;
  R[24] = (iss - iSp) / tiSp;
  R[47] = (Jrel_inf - Jrelnp) / tau_rel;
  R[48] = (Jrel_infp - Jrelp) / tau_relp;

  /* concentrations and membrane potential (:1099-1107) */
  const double cm = C[32], Acap = C[183], vmyo = C[184], vnsr = C[185], vjsr = C[186], vss = C[187];
  R[5] = (-((Ito + IKr + IKs + IK1 + IKb + Istim) - 2.0 * INaK) * cm * Acap) / (F * vmyo) + (JdiffK * vss) / vmyo;
  R[6] = (-ICaK * cm * Acap) / (F * vss) - JdiffK;
  R[3] = (-(INa + INaL + 3.0 * INaCa_i + 3.0 * INaK + INab) * Acap * cm) / (F * vmyo) + (JdiffNa * vss) / vmyo;
  R[4] = (-(ICaNa + 3.0 * INaCa_ss) * cm * Acap) / (F * vss) - JdiffNa;
  R[0] = -(INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa_i + INaCa_ss + INaK + INab + IKb + IpCa + ICab + Istim);
  R[2] = Bcass * (((-(ICaL - 2.0 * INaCa_ss) * cm * Acap) / (2.0 * F * vss) + (Jrel * vjsr) / vss) - Jdiff);
  R[9] = Bcai * ((((-((IpCa + ICab) - 2.0 * INaCa_i) * cm * Acap) / (2.0 * F * vmyo)) - (Jup * vnsr) / vmyo) + (Jdiff * vss) / vmyo);
  R[7] = Jup - (Jtr * vjsr) / vnsr;
  R[8] = Bcajsr * (Jtr - Jrel);
  /* accumulators (:1108-1110) */
  R[49] = INaL + ICaL + Ito + IKr + IKs + IK1;
  R[50] = INaL;
  R[51] = ICaL;

  if (A) {
    /* ALGEBRAIC slots read by the reference drivers (enum E_ALGEBRAIC_T) */
    A[0] = Istim; A[12] = vfrt; A[50] = ENa; A[53] = EK; A[54] = EKs;
    A[58] = INa; A[60] = INaL; A[66] = Ito; A[83] = ICaL; A[84] = ICaNa; A[87] = ICaK;
    A[90] = IKr; A[96] = IKs; A[98] = IK1; A[130] = INaCa_i; A[160] = INaCa_ss;
    A[179] = INaK; A[181] = IKb; A[185] = INab; A[189] = ICab; A[190] = IpCa;
    A[193] = Jrel; A[198] = Jup; A[42] = CaMKa; A[36] = CaMKb;
    A[44] = Bcai; A[46] = Bcass; A[48] = Bcajsr; A[191] = Jdiff; A[199] = Jtr;
  }
}
