// adams_host_check.cpp -- host build of the GPU Adams PECE path, for the CPU tests.
//
// Compiles the device integrator (2ddrugsimulation_amd/csrc/adams.hpp) and the device ORd 2011
// model (csrc/ord11_model.hpp) with g++ for the host, and runs them the way ord11_adams_kernel
// does: per launch, per cell, the scalars loaded from their SoA slots, `nout` de() calls to
// tout = t + dt, the scalars stored back.  Every difference-column, coefficient-slot and table
// index is checked (CEP_ADAMS_CHECK aborts), and tests/test_adams.py builds this file with
// AddressSanitizer and UndefinedBehaviorSanitizer, with and without fused multiply-adds.  The
// workspace arrays are allocated with exactly the sizes the library allocates, so an access
// past any of them is reported.
//
// Usage: adams_host_check <in> <out>, raw little-endian float64 files.
//   in : relerr, abserr, t_end, dt_out, every, ncell, then per cell conc, bcl, ic50[14], y0[41]
//        (y0[0] NaN: the model's initial states)
//   out: per sample (every `every` outputs) and cell: t, y[41]; then per cell: iflag, steps, rhs
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CEP_ADAMS_CHECK(cond)                                                            \
  do {                                                                                   \
    if (!(cond)) {                                                                       \
      fprintf(stderr, "adams index check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
      abort();                                                                           \
    }                                                                                    \
  } while (0)

// float intrinsics the device headers name in their FP32 paths, which this FP64 build never
// calls; the host HIP headers do not declare them
static inline float __builtin_amdgcn_rcpf(float a) { return 1.0f / a; }
static inline float __builtin_amdgcn_exp2f(float x) { return exp2f(x); }
static inline float __fdividef(float a, float b) { return a / b; }
static inline float __fsqrt_rn(float a) { return sqrtf(a); }

#include "../2ddrugsimulation_amd/csrc/adams.hpp"
#include "../2ddrugsimulation_amd/csrc/ord11_host.hpp"

namespace cep {
namespace ord {
MarkovTable g_markov;   // declared by the ORd-CiPA header; unused by ORd 2011
}
}  // namespace cep

using namespace cep;

struct PrmAcc {
  const double *p;
  int pitch;
  double operator()(int k) const { return p[(size_t)k * pitch]; }
};

int main(int argc, char **argv)
{
  if (argc != 3) {
    fprintf(stderr, "usage: adams_host_check in out\n");
    return 2;
  }
  std::vector<double> in;
  {
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    double x;
    while (fread(&x, sizeof(double), 1, f) == 1) in.push_back(x);
    fclose(f);
  }
  const double relerr = in[0], abserr = in[1], t_end = in[2], dt = in[3];
  const int every = (int)in[4], ncell = (int)in[5];
  const int NS = ord11::NS, rec = 2 + 14 + NS;
  if ((int)in.size() != 6 + ncell * rec || every < 1 || ncell < 1) {
    fprintf(stderr, "bad input size\n");
    return 2;
  }
  const int P = (ncell + 63) / 64 * 64;          // the library's pitch
  // the library's allocations (cep_batch: d_y, d_prm, and reset_integrators' Adams workspace)
  std::vector<double> y((size_t)P * NS), prm((size_t)P * ord11::NP);
  std::vector<double> phi((size_t)P * adams::COLS * NS), vec((size_t)P * 5 * NS);
  std::vector<double> coef((size_t)P * adams::C_N), sc((size_t)P * adams::S_N);
  std::vector<int> si((size_t)P * adams::I_N), status((size_t)P, 0);
  std::vector<long long> stats((size_t)2 * P, 0);
  for (int c = 0; c < ncell; ++c) {
    const double *r = &in[6 + (size_t)c * rec];
    host::O11CellIn ci{};
    ci.conc = r[0];
    ci.bcl = r[1];
    for (int k = 0; k < 14; ++k) ci.ic50[k] = r[2 + k];
    double p[ord11::NP];
    host::o11_params(ci, false, 1.0, p);
    for (int k = 0; k < ord11::NP; ++k) prm[(size_t)k * P + c] = p[k];
    const double *y0 = r + 16;
    for (int s = 0; s < NS; ++s) y[(size_t)s * P + c] = std::isnan(y0[0]) ? host::kO11Init[s] : y0[s];
  }
  adams::Buffers A;
  A.ncell = ncell; A.pitch = P; A.phi = phi.data(); A.vec = vec.data(); A.coef = coef.data();
  A.sc = sc.data(); A.si = si.data(); A.stats = stats.data();

  FILE *fo = fopen(argv[2], "wb");
  if (!fo) { perror(argv[2]); return 2; }
  const long imax = std::lround(t_end / dt);
  std::vector<double> tcell(ncell, 0.0);
  std::vector<int> flag(ncell, 0);
  for (long done = 0; done < imax; done += every) {
    const int nout = (int)std::min<long>(every, imax - done);
    // one launch of ord11_adams_kernel: every cell (lane) advances nout output points
    for (int c = 0; c < ncell; ++c) {
      if (status[c] != 0) continue;
      const PrmAcc pa{ prm.data() + c, P };
      auto f = [&](double t, const double *yy, double *yp) {
        ord11::Gates g;
        ord11::Currents cur;
        ord11::rhs<double>(t, yy, yp, pa, cur, g);
      };
      adams::Lane<ord11::NS, decltype(f)> L{ A, (size_t)c, (size_t)P, f };
      L.load();
      double yl[ord11::NS];
      for (int i = 0; i < NS; ++i) yl[i] = y[(size_t)i * P + c];
      double t = sc[(size_t)adams::S_T * P + c];
      int iflag = si[(size_t)adams::I_IFLAG * P + c];
      for (int o = 0; o < nout; ++o) {
        const double tout = t + dt;
        iflag = L.de(yl, t, tout, relerr, abserr, iflag == 0 ? 1 : iflag);
        if (iflag != 2) { status[c] = 1; break; }
      }
      for (int i = 0; i < NS; ++i) y[(size_t)i * P + c] = yl[i];
      sc[(size_t)adams::S_T * P + c] = t;
      si[(size_t)adams::I_IFLAG * P + c] = iflag;
      L.store();
      tcell[c] = t;
      flag[c] = iflag;
    }
    if (nout == every)
      for (int c = 0; c < ncell; ++c) {
        fwrite(&tcell[c], sizeof(double), 1, fo);
        for (int i = 0; i < NS; ++i) fwrite(&y[(size_t)i * P + c], sizeof(double), 1, fo);
      }
  }
  for (int c = 0; c < ncell; ++c) {
    const double tr[3] = { (double)flag[c], (double)stats[c], (double)stats[(size_t)P + c] };
    fwrite(tr, sizeof(double), 3, fo);
  }
  fclose(fo);
  return 0;
}
