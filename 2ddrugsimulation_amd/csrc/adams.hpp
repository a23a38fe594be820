// adams.hpp -- Shampine-Gordon variable-order Adams PECE (DE / STEP / INTRP), one cell per lane.
//
// The integrator of libs/ode.cpp (ode :589-805, de :14-403, step :970-1645, intrp :446-586) as
// CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:119-146 drives it: repeated calls to
// tout = t + dt that continue the integration (iflag 2), the solver stepping past tout and
// interpolating back.  Per cell the state lives in HBM as SoA (element j of cell c at
// j * pitch + c): the 16 difference columns phi, the vectors yy / wt / p / yp / ypout, the
// coefficient arrays psi / alpha / beta / sigma / v / w / g and the scalars; a launch keeps
// the scalars in registers and streams the vectors.  The arithmetic follows the published
// algorithm step by step (the same order as the oracle, which is bit-identical to ode.cpp).
#pragma once
#include "common.hpp"

// Index check of every difference column, coefficient slot and table entry.  No-op on the
// device; a host build of this header (tests/adams_host_check.cpp) defines it to abort, and runs
// the integrator under AddressSanitizer.
#ifndef CEP_ADAMS_CHECK
#define CEP_ADAMS_CHECK(cond) ((void)0)
#endif

namespace cep {
namespace adams {

constexpr int COLS = 16;
constexpr int MAXNUM = 500;
constexpr double UROUND = 2.220446049250313E-016;
// coefficient slots
enum : int { C_PSI = 0, C_ALPHA = 12, C_BETA = 24, C_SIG = 36, C_V = 49, C_W = 61, C_G = 73, C_N = 86 };
// scalar slots
enum : int { S_X = 0, S_H, S_HOLD, S_TOLD, S_DELSGN, S_T, S_N };
enum : int { I_NS = 0, I_K, I_KOLD, I_ISNOLD, I_START, I_PHASE1, I_NORND, I_IFLAG, I_N };

struct Buffers {
  int ncell, pitch;
  double *phi;      // [COLS * N]
  double *vec;      // [5 * N]: yy, wt, p, yp, ypout
  double *coef;     // [C_N]
  double *sc;       // [S_N]
  int *si;          // [I_N]
  long long *stats; // [2]: steps, right-hand sides
};

__device__ __constant__ static const double kGstr[13] = { 0.50, 0.0833, 0.0417, 0.0264, 0.0188, 0.0143, 0.0114,
                                                          0.00936, 0.00789, 0.00679, 0.00592, 0.00524, 0.00468 };
__device__ __constant__ static const double kTwo[13] = { 2.0, 4.0, 8.0, 16.0, 32.0, 64.0, 128.0, 256.0, 512.0,
                                                         1024.0, 2048.0, 4096.0, 8192.0 };

CEP_D double sgn(double a) { return a < 0.0 ? -1.0 : 1.0; }
CEP_D double dmax(double a, double b) { return b < a ? a : b; }
CEP_D double dmin(double a, double b) { return b < a ? b : a; }
CEP_D double sq(double a) { return a * a; }

// One lane's integrator.  F: callable rhs(t, const double *y_in_registers, double *yp_out).
template <int N, typename F>
struct Lane {
  const Buffers &B;
  size_t c, P;
  F &f;
  // scalars
  double x, h, hold, told, delsgn;
  int ns, k, kold, isnold, start, phase1, nornd;
  long long nsteps = 0, nfe = 0;

  CEP_D double &PHI(int l, int col) const
  {
    CEP_ADAMS_CHECK(l >= 0 && l < N && col >= 0 && col < COLS);
    return B.phi[((size_t)col * N + l) * P + c];
  }
  CEP_D double &VEC(int which, int l) const
  {
    CEP_ADAMS_CHECK(l >= 0 && l < N);
    return B.vec[((size_t)which * N + l) * P + c];
  }
  CEP_D double &YY(int l) const { return VEC(0, l); }
  CEP_D double &WT(int l) const { return VEC(1, l); }
  CEP_D double &PP(int l) const { return VEC(2, l); }
  CEP_D double &YP(int l) const { return VEC(3, l); }
  CEP_D double &YPOUT(int l) const { return VEC(4, l); }
  // slot i of the coefficient array that starts at slot `base` and holds `len` entries
  CEP_D double &CO(int base, int len, int i) const
  {
    (void)len;   // read by the host build's index check only
    CEP_ADAMS_CHECK(i >= 0 && i < len);
    return B.coef[(size_t)(base + i) * P + c];
  }
  CEP_D double &psi(int i) const { return CO(C_PSI, C_ALPHA - C_PSI, i); }
  CEP_D double &alpha(int i) const { return CO(C_ALPHA, C_BETA - C_ALPHA, i); }
  CEP_D double &beta(int i) const { return CO(C_BETA, C_SIG - C_BETA, i); }
  CEP_D double &sig(int i) const { return CO(C_SIG, C_V - C_SIG, i); }
  CEP_D double &v(int i) const { return CO(C_V, C_W - C_V, i); }
  CEP_D double &w(int i) const { return CO(C_W, C_G - C_W, i); }
  CEP_D double &g(int i) const { return CO(C_G, C_N - C_G, i); }
  CEP_D static double gstr(int i) { CEP_ADAMS_CHECK(i >= 0 && i < 13); return kGstr[i]; }
  CEP_D static double two(int i) { CEP_ADAMS_CHECK(i >= 0 && i < 13); return kTwo[i]; }

  CEP_D void load()
  {
    x = B.sc[S_X * P + c]; h = B.sc[S_H * P + c]; hold = B.sc[S_HOLD * P + c];
    told = B.sc[S_TOLD * P + c]; delsgn = B.sc[S_DELSGN * P + c];
    ns = B.si[I_NS * P + c]; k = B.si[I_K * P + c]; kold = B.si[I_KOLD * P + c];
    isnold = B.si[I_ISNOLD * P + c]; start = B.si[I_START * P + c]; phase1 = B.si[I_PHASE1 * P + c];
    nornd = B.si[I_NORND * P + c];
  }
  CEP_D void store() const
  {
    B.sc[S_X * P + c] = x; B.sc[S_H * P + c] = h; B.sc[S_HOLD * P + c] = hold;
    B.sc[S_TOLD * P + c] = told; B.sc[S_DELSGN * P + c] = delsgn;
    B.si[I_NS * P + c] = ns; B.si[I_K * P + c] = k; B.si[I_KOLD * P + c] = kold;
    B.si[I_ISNOLD * P + c] = isnold; B.si[I_START * P + c] = start; B.si[I_PHASE1 * P + c] = phase1;
    B.si[I_NORND * P + c] = nornd;
    B.stats[0 * P + c] += nsteps;
    B.stats[1 * P + c] += nfe;
  }

  // right-hand side of a vector given by an accessor (YY / PP) into YP
  template <typename A>
  CEP_D void rhs_into_yp(double t, A src)
  {
    double y[N], yp[N];
#pragma unroll
    for (int l = 0; l < N; ++l) y[l] = src(l);
    f(t, y, yp);
#pragma unroll
    for (int l = 0; l < N; ++l) YP(l) = yp[l];
    ++nfe;
  }

  // INTRP at xout into yout (registers)
  CEP_D void interpolate(double xout, double *yout)
  {
    double gi[13], rho[13], wi[13];
    const double hi = xout - x;
    const int ki = kold + 1;
    CEP_ADAMS_CHECK(ki >= 1 && ki <= 13);
    for (int i = 1; i <= ki; ++i) wi[i - 1] = 1.0 / (double)i;
    gi[0] = 1.0;
    rho[0] = 1.0;
    double term = 0.0;
    for (int j = 2; j <= ki; ++j) {
      const double psijm1 = psi(j - 2);
      const double gamma = (hi + term) / psijm1;
      const double eta = hi / psijm1;
      for (int i = 1; i <= ki + 1 - j; ++i) wi[i - 1] = gamma * wi[i - 1] - eta * wi[i];
      gi[j - 1] = wi[0];
      rho[j - 1] = gamma * rho[j - 2];
      term = psijm1;
    }
    double yo[N], ypo[N];
#pragma unroll
    for (int l = 0; l < N; ++l) { yo[l] = 0.0; ypo[l] = 0.0; }
    for (int j = 1; j <= ki; ++j) {
      const int i = ki + 1 - j;
      const double gg = gi[i - 1], rr = rho[i - 1];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        const double ph = PHI(l, i - 1);
        yo[l] = yo[l] + gg * ph;
        ypo[l] = ypo[l] + rr * ph;
      }
    }
#pragma unroll
    for (int l = 0; l < N; ++l) {
      YPOUT(l) = ypo[l];
      yout[l] = YY(l) + hi * yo[l];
    }
  }

  // STEP; returns true on a crash (h too small for x, or eps below round-off)
  CEP_D bool step(double &eps)
  {
    const double twou = 2.0 * UROUND, fouru = 2.0 * twou;
    double erk, erkm1 = 0.0, erkm2, err, absh, temp1, temp2;
    int knew, kp1, kp2, km1, km2;
    if (fabs(h) < fouru * fabs(x)) {
      h = fouru * fabs(x) * sgn(h);
      return true;
    }
    const double p5eps = 0.5 * eps;
    double round = 0.0;
    for (int l = 0; l < N; ++l) round = round + sq(YY(l) / WT(l));
    round = twou * sqrt(round);
    if (p5eps < round) {
      eps = 2.0 * round * (1.0 + fouru);
      return true;
    }
    g(0) = 1.0;
    g(1) = 0.5;
    sig(0) = 1.0;
    if (start) {
      rhs_into_yp(x, [&](int l) { return YY(l); });
      for (int l = 0; l < N; ++l) { PHI(l, 0) = YP(l); PHI(l, 1) = 0.0; }
      double total = 0.0;
      for (int l = 0; l < N; ++l) total = total + sq(YP(l) / WT(l));
      total = sqrt(total);
      absh = fabs(h);
      if (eps < 16.0 * total * h * h) absh = 0.25 * sqrt(eps / total);
      h = dmax(absh, fouru * fabs(x)) * sgn(h);
      hold = 0.0;
      k = 1;
      kold = 0;
      start = 0;
      phase1 = 1;
      nornd = 1;
      if (p5eps <= 100.0 * round) {
        nornd = 0;
        for (int l = 0; l < N; ++l) PHI(l, 14) = 0.0;
      }
    }
    int ifail = 0;
    for (;;) {
      kp1 = k + 1; kp2 = k + 2; km1 = k - 1; km2 = k - 2;
      if (h != hold) ns = 0;
      if (ns <= kold) ns = ns + 1;
      const int nsp1 = ns + 1;
      if (ns <= k) {
        beta(ns - 1) = 1.0;
        alpha(ns - 1) = 1.0 / (double)ns;
        temp1 = h * (double)ns;
        sig(nsp1 - 1) = 1.0;
        for (int i = nsp1; i <= k; ++i) {
          temp2 = psi(i - 2);
          psi(i - 2) = temp1;
          beta(i - 1) = beta(i - 2) * psi(i - 2) / temp2;
          temp1 = temp2 + h;
          alpha(i - 1) = h / temp1;
          sig(i) = (double)i * alpha(i - 1) * sig(i - 1);
        }
        psi(k - 1) = temp1;
        if (ns <= 1) {
          for (int iq = 1; iq <= k; ++iq) {
            v(iq - 1) = 1.0 / (double)(iq * (iq + 1));
            w(iq - 1) = v(iq - 1);
          }
        } else {
          if (kold < k) {
            v(k - 1) = 1.0 / (double)(k * kp1);
            for (int j = 1; j <= ns - 2; ++j) {
              const int i = k - j;
              v(i - 1) = v(i - 1) - alpha(j) * v(i);
            }
          }
          for (int iq = 1; iq <= kp1 - ns; ++iq) {
            v(iq - 1) = v(iq - 1) - alpha(ns - 1) * v(iq);
            w(iq - 1) = v(iq - 1);
          }
          g(nsp1 - 1) = w(0);
        }
        for (int i = ns + 2; i <= kp1; ++i) {
          for (int iq = 1; iq <= kp2 - i; ++iq) w(iq - 1) = w(iq - 1) - alpha(i - 2) * w(iq);
          g(i - 1) = w(0);
        }
      }
      for (int i = nsp1; i <= k; ++i) {
        const double bt = beta(i - 1);
        for (int l = 0; l < N; ++l) PHI(l, i - 1) = bt * PHI(l, i - 1);
      }
      double p[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        PHI(l, kp2 - 1) = PHI(l, kp1 - 1);
        PHI(l, kp1 - 1) = 0.0;
        p[l] = 0.0;
      }
      for (int j = 1; j <= k; ++j) {
        const int i = kp1 - j;
        const double gg = g(i - 1);
#pragma unroll
        for (int l = 0; l < N; ++l) {
          const double ph = PHI(l, i - 1);
          p[l] = p[l] + gg * ph;
          PHI(l, i - 1) = ph + PHI(l, i);
        }
      }
      if (!nornd) {
#pragma unroll
        for (int l = 0; l < N; ++l) {
          const double yl = YY(l);
          const double tau = h * p[l] - PHI(l, 14);
          p[l] = yl + tau;
          PHI(l, 15) = (p[l] - yl) - tau;
        }
      } else {
#pragma unroll
        for (int l = 0; l < N; ++l) p[l] = YY(l) + h * p[l];
      }
#pragma unroll
      for (int l = 0; l < N; ++l) PP(l) = p[l];
      const double xold = x;
      x = x + h;
      absh = fabs(h);
      {
        double yp[N];
        f(x, p, yp);
        ++nfe;
#pragma unroll
        for (int l = 0; l < N; ++l) YP(l) = yp[l];
      }
      erkm2 = 0.0; erkm1 = 0.0; erk = 0.0;
      for (int l = 0; l < N; ++l) {
        const double ypl = YP(l), ph0 = PHI(l, 0), wl = WT(l);
        if (0 < km2) erkm2 = erkm2 + sq((PHI(l, km1 - 1) + ypl - ph0) / wl);
        if (0 <= km2) erkm1 = erkm1 + sq((PHI(l, k - 1) + ypl - ph0) / wl);
        erk = erk + sq((ypl - ph0) / wl);
      }
      if (0 < km2) erkm2 = absh * sig(km1 - 1) * gstr(km2 - 1) * sqrt(erkm2);
      if (0 <= km2) erkm1 = absh * sig(k - 1) * gstr(km1 - 1) * sqrt(erkm1);
      err = absh * sqrt(erk) * (g(k - 1) - g(kp1 - 1));
      erk = absh * sqrt(erk) * sig(kp1 - 1) * gstr(k - 1);
      knew = k;
      if (0 < km2) {
        if (dmax(erkm1, erkm2) <= erk) knew = km1;
      } else if (0 == km2) {
        if (erkm1 <= 0.5 * erk) knew = km1;
      }
      if (err <= eps) break;
      phase1 = 0;
      x = xold;
      for (int i = 1; i <= k; ++i) {
        const double bt = beta(i - 1);
        for (int l = 0; l < N; ++l) PHI(l, i - 1) = (PHI(l, i - 1) - PHI(l, i)) / bt;
      }
      for (int i = 2; i <= k; ++i) psi(i - 2) = psi(i - 1) - h;
      ifail = ifail + 1;
      temp2 = 0.5;
      if (3 < ifail && p5eps < 0.25 * erk) temp2 = sqrt(p5eps / erk);
      if (3 <= ifail) knew = 1;
      h = temp2 * h;
      k = knew;
      // a NaN error estimate fails every attempt: with h NaN, or halved to 0 at x = 0, the
      // published test alone never ends the loop, so those two cases crash here as well (finite
      // non-zero h: the published test unchanged)
      if (fabs(h) < fouru * fabs(x) || !(fabs(h) > 0.0)) {
        h = fabs(fouru * fabs(x)) * sgn(h);
        eps = eps + eps;
        return true;
      }
    }
    kold = k;
    hold = h;
    const double gk = g(kp1 - 1);
    double y[N];
    if (!nornd) {
#pragma unroll
      for (int l = 0; l < N; ++l) {
        const double pl = PP(l);
        const double rho = h * gk * (YP(l) - PHI(l, 0)) - PHI(l, 15);
        y[l] = pl + rho;
        PHI(l, 14) = (y[l] - pl) - rho;
      }
    } else {
#pragma unroll
      for (int l = 0; l < N; ++l) y[l] = PP(l) + h * gk * (YP(l) - PHI(l, 0));
    }
#pragma unroll
    for (int l = 0; l < N; ++l) YY(l) = y[l];
    {
      double yp[N];
      f(x, y, yp);
      ++nfe;
#pragma unroll
      for (int l = 0; l < N; ++l) YP(l) = yp[l];
    }
    for (int l = 0; l < N; ++l) {
      const double d = YP(l) - PHI(l, 0);
      PHI(l, kp1 - 1) = d;
      PHI(l, kp2 - 1) = d - PHI(l, kp2 - 1);
    }
    for (int i = 1; i <= k; ++i)
      for (int l = 0; l < N; ++l) PHI(l, i - 1) = PHI(l, i - 1) + PHI(l, kp1 - 1);
    double erkp1 = 0.0;
    if (knew == km1 || k == 12) phase1 = 0;
    if (phase1) {
      k = kp1;
      erk = erkp1;
    } else if (knew == km1) {
      k = km1;
      erk = erkm1;
    } else if (kp1 <= ns) {
      for (int l = 0; l < N; ++l) erkp1 = erkp1 + sq(PHI(l, kp2 - 1) / WT(l));
      erkp1 = absh * gstr(kp1 - 1) * sqrt(erkp1);
      if (k == 1) {
        if (erkp1 < 0.5 * erk) { k = kp1; erk = erkp1; }
      } else if (erkm1 <= dmin(erk, erkp1)) {
        k = km1; erk = erkm1;
      } else if (erkp1 < erk && k < 12) {
        k = kp1; erk = erkp1;
      }
    }
    double hnew = h + h;
    if (!phase1 && p5eps < erk * two(k)) {
      hnew = h;
      if (p5eps < erk) {
        temp2 = (double)(k + 1);
        const double r = pow(p5eps / erk, 1.0 / temp2);
        hnew = absh * dmax(0.5, dmin(0.9, r));
        hnew = fabs(dmax(hnew, fouru * fabs(x))) * sgn(h);
      }
    }
    h = hnew;
    ++nsteps;
    return false;
  }

  // DE for one output: from t (== told on continuation) to tout; y in registers in and out.
  // Returns the iflag of the reference (2 ok, 3 crash, 4 / 5 more than MAXNUM steps, 6 input).
  CEP_D int de(double *y, double &t, double tout, double relerr, double abserr, int iflag)
  {
    const double fouru = 4.0 * UROUND;
    if (t == tout || relerr < 0.0 || abserr < 0.0) return 6;
    double eps = dmax(relerr, abserr);
    if (eps <= 0.0 || iflag == 0) return 6;
    const int isn = iflag < 0 ? -1 : 1;
    iflag = iflag < 0 ? -iflag : iflag;
    if (iflag != 1 && (t != told || iflag < 2 || 5 < iflag)) return 6;
    const double del = tout - t;
    const double absdel = fabs(del);
    const double tend = isn < 0 ? tout : t + 10.0 * del;
    int nostep = 0, kle4 = 0, stiff = 0;
    const double releps = relerr / eps, abseps = abserr / eps;
    if (iflag == 1 || isnold < 0 || delsgn * del <= 0.0) {
      start = 1;
      x = t;
#pragma unroll
      for (int l = 0; l < N; ++l) YY(l) = y[l];
      delsgn = sgn(del);
      h = dmax(fabs(tout - x), fouru * fabs(x)) * sgn(tout - x);
    }
    for (;;) {
      if (absdel <= fabs(x - t)) {
        interpolate(tout, y);
        t = tout;
        told = t;
        isnold = isn;
        return 2;
      }
      if (isn <= 0 && fabs(tout - x) < fouru * fabs(x)) {
        h = tout - x;
        rhs_into_yp(x, [&](int l) { return YY(l); });
#pragma unroll
        for (int l = 0; l < N; ++l) y[l] = YY(l) + h * YP(l);
        t = tout;
        told = t;
        isnold = isn;
        return 2;
      }
      if (MAXNUM <= nostep) {
#pragma unroll
        for (int l = 0; l < N; ++l) y[l] = YY(l);
        t = x;
        told = t;
        isnold = 1;
        return isn * (stiff ? 5 : 4);
      }
      h = dmin(fabs(h), fabs(tend - x)) * sgn(h);
      for (int l = 0; l < N; ++l) WT(l) = releps * fabs(YY(l)) + abseps;
      if (step(eps)) {
#pragma unroll
        for (int l = 0; l < N; ++l) y[l] = YY(l);
        t = x;
        told = t;
        isnold = 1;
        return isn * 3;
      }
      nostep = nostep + 1;
      kle4 = kle4 + 1;
      if (4 < kold) kle4 = 0;
      if (50 <= kle4) stiff = 1;
    }
  }
};

}  // namespace adams
}  // namespace cep
