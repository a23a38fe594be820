// ref_golden.cpp -- ORACLE helper (test infrastructure only).
//
// Links the reference's own CellML model sources, compiled from where they lie under
// the reference tree by oracle/Makefile, and evaluates them on inputs drawn by
// tests/golden/make_golden.py.  Its outputs become the golden vectors that pin the
// C restatement in oracle/*.c and the HIP kernels.  Nothing here is shipped.
//
// Modes (all files are raw little-endian float64):
//   ord_init  in: [celltype, conc, herg[19]] per record            out: C[206] S[52]
//   ord_rhs   in: [celltype, conc, herg[19], t, S[52]] per record  out: R[52] A[200]
//   ord_rk4   in: [celltype, conc, herg[19], t_end, h, every]      out: S[52] per sample
//   tt_rhs    in: [celltype, t, S[19]] per record                  out: R[19] A[70]
//   tt_rl     in: [celltype, t_end, dt, every]                      out: S[19] per sample
//   tt_init   in: [celltype] per record                             out: C[53] S[19]
//   o11_init  in: [celltype, conc, hill[14], dutta] per record      out: C[142] S[44]
//   o11_rhs   in: [t, C[142], S[41]] per record                     out: R[41] A[198]
// The ORd 2011 modes use the reference's patch-clamp copy of the model
// (CEPPatch_CiPA_Sim/patches/Ohara_Rudy_2011.cpp), whose state and constant layout is that of
// DrugSimulation's (states 0-40; it adds three charge accumulators at 41-43; constant 140 is
// the stimulus period).  Its initConsts leaves PKNa and KmCaAct unset and handles the cell
// type and Dutta's scaling differently from DrugSimulation's copy, so o11_rhs takes the
// constants as input and evaluates computeRates alone.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <string>
// The reference's enum headers clash when two models share a translation unit, so this file
// is compiled once per model: -DREF_ORD (ORd-CiPA), -DREF_TT (TT06), -DREF_O11 (ORd 2011).
#if defined(REF_ORD)
#include "ohara_rudy_cipa_v1_2017.hpp"
#elif defined(REF_O11)
#include "Ohara_Rudy_2011.hpp"
#else
#include "ten_tusscher_model_2006_IK1Ko_endo_units.hpp"
#include "ten_tusscher_model_2006_IK1Ko_epi_units.hpp"
#include "ten_tusscher_model_2006_IK1Ko_M_units.hpp"
#endif

static std::vector<double> read_all(const char *path)
{
  std::vector<double> v;
  FILE *f = fopen(path, "rb");
  if (!f) { perror(path); exit(2); }
  double x;
  while (fread(&x, sizeof(double), 1, f) == 1) v.push_back(x);
  fclose(f);
  return v;
}

static void put(FILE *f, const double *x, size_t n) { fwrite(x, sizeof(double), n, f); }

#if defined(REF_TT)
static CellML *make_tt(int ct)
{
  CellML *c;
  if (ct == 1) c = new ten_tusscher_model_2006_IK1Ko_epi_units();
  else if (ct == 2) c = new ten_tusscher_model_2006_IK1Ko_M_units();
  else c = new ten_tusscher_model_2006_IK1Ko_endo_units();
  c->mutation = 0; c->isMutated = 0; c->isS1 = 0; c->isEctopic = 0;
  c->initConsts();
  return c;
}

#endif
#if defined(REF_O11)
static Ohara_Rudy_2011 *make_o11(const double *rec)
{
  Ohara_Rudy_2011 *c = new Ohara_Rudy_2011();
  double hill[14];
  memcpy(hill, rec + 2, sizeof(hill));
  c->initConsts(rec[0], rec[1], hill, rec[16] != 0.0);
  return c;
}
#endif
#if defined(REF_ORD)
static ohara_rudy_cipa_v1_2017 *make_ord(const double *rec)
{
  ohara_rudy_cipa_v1_2017 *c = new ohara_rudy_cipa_v1_2017();
  c->mutation = 0; c->isMutated = 0; c->isS1 = 0; c->isEctopic = 0;
  double herg[19];
  memcpy(herg, rec + 2, sizeof(herg));
  c->initConsts(rec[0], rec[1], herg, "golden");
  return c;
}
#endif

int main(int argc, char **argv)
{
  if (argc != 4) { fprintf(stderr, "usage: %s mode in out\n", argv[0]); return 2; }
  std::string mode = argv[1];
  std::vector<double> in = read_all(argv[2]);
  FILE *out = fopen(argv[3], "wb");
  if (!out) { perror(argv[3]); return 2; }
  if (0) {
#if defined(REF_ORD)
  } else if (mode == "ord_init") {
    for (size_t k = 0; k + 21 <= in.size(); k += 21) {
      ohara_rudy_cipa_v1_2017 *c = make_ord(&in[k]);
      put(out, c->CONSTANTS, 206); put(out, c->STATES, 52);
      delete c;
    }
  } else if (mode == "ord_rhs") {
    const size_t rl = 21 + 1 + 52;
    for (size_t k = 0; k + rl <= in.size(); k += rl) {
      ohara_rudy_cipa_v1_2017 *c = make_ord(&in[k]);
      double S[52], R[52], A[200];
      memcpy(S, &in[k + 22], sizeof(S));
      memset(R, 0, sizeof(R)); memset(A, 0, sizeof(A));
      c->computeRates(in[k + 21], c->CONSTANTS, R, S, A);
      put(out, R, 52); put(out, A, 200);
      delete c;
    }
  } else if (mode == "ord_rk4") {
    // classical RK4 on the reference RHS at a small fixed step: a textbook integrator
    // applied to the reference model, used to pin the BDF trajectory
    ohara_rudy_cipa_v1_2017 *c = make_ord(&in[0]);
    const double t_end = in[21], h = in[22];
    const long every = (long)in[23];
    const int n = 52;
    double y[52], k1[52], k2[52], k3[52], k4[52], yt[52], A[200];
    memcpy(y, c->STATES, sizeof(y));
    const long nsteps = (long)(t_end / h + 0.5);
    put(out, y, n);
    for (long s = 0; s < nsteps; ++s) {
      const double t = s * h;
      c->computeRates(t, c->CONSTANTS, k1, y, A);
      for (int i = 0; i < n; ++i) yt[i] = y[i] + 0.5 * h * k1[i];
      c->computeRates(t + 0.5 * h, c->CONSTANTS, k2, yt, A);
      for (int i = 0; i < n; ++i) yt[i] = y[i] + 0.5 * h * k2[i];
      c->computeRates(t + 0.5 * h, c->CONSTANTS, k3, yt, A);
      for (int i = 0; i < n; ++i) yt[i] = y[i] + h * k3[i];
      c->computeRates(t + h, c->CONSTANTS, k4, yt, A);
      for (int i = 0; i < n; ++i) y[i] += h / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
      if ((s + 1) % every == 0) put(out, y, n);
    }
#endif
#if defined(REF_O11)
  } else if (mode == "o11_init") {
    for (size_t k = 0; k + 17 <= in.size(); k += 17) {
      Ohara_Rudy_2011 *c = make_o11(&in[k]);
      put(out, c->CONSTANTS, 142); put(out, c->STATES, 44);
      delete c;
    }
  } else if (mode == "o11_rhs") {
    const size_t rl = 1 + 142 + 41;
    Ohara_Rudy_2011 *c = new Ohara_Rudy_2011();
    for (size_t k = 0; k + rl <= in.size(); k += rl) {
      double Cc[142], S[44], R[44], A[198];
      memcpy(Cc, &in[k + 1], sizeof(Cc));
      memset(S, 0, sizeof(S));
      memcpy(S, &in[k + 143], 41 * sizeof(double));
      memset(R, 0, sizeof(R)); memset(A, 0, sizeof(A));
      // twice on the same inputs: the copy reads ALGEBRAIC[P] in its INaK block
      // (Ohara_Rudy_2011.cpp:1459) before it sets it (:1464); the second call sees this P
      c->computeRates(in[k], Cc, R, S, A);
      c->computeRates(in[k], Cc, R, S, A);
      put(out, R, 41); put(out, A, 198);
    }
    delete c;
#endif
#if defined(REF_TT)
  } else if (mode == "tt_init") {
    for (size_t k = 0; k < in.size(); ++k) {
      CellML *c = make_tt((int)in[k]);
      put(out, c->CONSTANTS, 53); put(out, c->STATES, 19);
      delete c;
    }
  } else if (mode == "tt_rhs") {
    const size_t rl = 2 + 19;
    for (size_t k = 0; k + rl <= in.size(); k += rl) {
      CellML *c = make_tt((int)in[k]);
      double S[19], R[19], A[70];
      memcpy(S, &in[k + 2], sizeof(S));
      memset(R, 0, sizeof(R)); memset(A, 0, sizeof(A));
      c->computeRates(in[k + 1], c->CONSTANTS, R, S, A);
      put(out, R, 19); put(out, A, 70);
      delete c;
    }
  } else if (mode == "tt_rl") {
    CellML *c = make_tt((int)in[0]);
    const double t_end = in[1], dt = in[2];
    const long every = (long)in[3];
    const long nsteps = (long)(t_end / dt + 0.5);
    put(out, c->STATES, 19);
    // time advances by accumulation, as CEP7Cell/modules/population_tn2006.cpp:143-153
    double tcurr = 0.0;
    for (long s = 0; s < nsteps; ++s) {
      c->computeRates(tcurr, c->CONSTANTS, c->RATES, c->STATES, c->ALGEBRAIC);
      c->solveAnalytical(dt);
      tcurr += dt;
      if ((s + 1) % every == 0) put(out, c->STATES, 19);
    }
    delete c;
#endif
  } else {
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
  }
  fclose(out);
  return 0;
}
