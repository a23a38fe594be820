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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <string>
// The reference's enum headers clash when both models share a translation unit, so
// this file is compiled twice: -DREF_ORD (ORd-CiPA modes) and -DREF_TT (TT06 modes).
#if defined(REF_ORD)
#include "ohara_rudy_cipa_v1_2017.hpp"
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
#if defined(REF_TT)
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
