// ref_ode.cpp -- ORACLE helper (test infrastructure only).
//
// Drives the reference's own Adams PECE solver (libs/ode.cpp, compiled from where it lies by
// oracle/Makefile) the way CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:119-146 does:
// repeated ode() calls to tout = dt, 2 dt, ..., continuing with iflag = 2.  Its outputs become
// the golden vectors that pin oracle/adams.c and the HIP Adams kernel.  Nothing here is shipped.
//
// Usage: ref_ode <mode> <in> <out>; all files raw little-endian float64.
//   in : relerr, abserr, t_end, dt_out, every, then for mode "ord11": conc, bcl, ic50[14]
//   out: per sample (every `every` outputs): t, y[neqn]; then a trailer: iflag, n_outputs
// Modes: linear (3 equations, y' = A y), vdp (Van der Pol, mu = 1), ord11 (the oracle's ORd 2011
// right-hand side, oracle/ord2011.c, single-cell stimulus).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ode.hpp"

extern "C" {
#include "oracle.h"
}

static double g_C[ORD11_NC];

static void f_linear(double, double y[], double yp[])
{
  yp[0] = -0.1 * y[0] + 1.0 * y[1];
  yp[1] = -1.0 * y[0] - 0.1 * y[1];
  yp[2] = -0.5 * y[2];
}

static void f_vdp(double, double y[], double yp[])
{
  yp[0] = y[1];
  yp[1] = (1.0 - y[0] * y[0]) * y[1] - y[0];
}

static void f_ord11(double t, double y[], double yp[])
{
  double A[ORD11_NA];
  ord11_rates(t, g_C, yp, y, A);
}

int main(int argc, char **argv)
{
  if (argc != 4) { fprintf(stderr, "usage: ref_ode mode in out\n"); return 2; }
  const std::string mode = argv[1];
  std::vector<double> in;
  {
    FILE *f = fopen(argv[2], "rb");
    if (!f) { perror(argv[2]); return 2; }
    double x;
    while (fread(&x, sizeof(double), 1, f) == 1) in.push_back(x);
    fclose(f);
  }
  const double relerr = in[0], abserr = in[1], t_end = in[2], dt = in[3];
  const int every = (int)in[4];
  int n = 0;
  void (*f)(double, double[], double[]) = nullptr;
  std::vector<double> y;
  if (mode == "linear") { n = 3; f = f_linear; y = { 1.0, 0.0, 2.0 }; }
  else if (mode == "vdp") { n = 2; f = f_vdp; y = { 2.0, 0.0 }; }
  else if (mode == "ord11") {
    n = ORD11_NS;
    f = f_ord11;
    y.resize(n);
    ord11_init_consts(g_C, y.data(), 0.0);
    ord11_apply_drug(g_C, in[5], &in[7]);
    g_C[140] = in[6];   // BCL
  } else {
    fprintf(stderr, "unknown mode\n");
    return 2;
  }
  std::vector<double> work(100 + 21 * n);
  int iwork[5] = { 0, 0, 0, 0, 0 };
  FILE *fo = fopen(argv[3], "wb");
  double t = 0.0, tnext = dt;
  int iflag = 1;
  const long imax = (long)std::llround(t_end / dt);
  long iout = 0;
  while (iout < imax) {
    ode(f, n, y.data(), t, tnext, relerr, abserr, iflag, work.data(), iwork);
    if (iflag != 2) break;
    iout++;
    tnext += dt;
    if (iout % every == 0) {
      fwrite(&t, sizeof(double), 1, fo);
      fwrite(y.data(), sizeof(double), n, fo);
    }
  }
  const double trailer[2] = { (double)iflag, (double)iout };
  fwrite(trailer, sizeof(double), 2, fo);
  fclose(fo);
  return 0;
}
