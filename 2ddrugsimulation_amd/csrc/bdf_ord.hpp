// bdf_ord.hpp -- per-cell CVODE-style BDF for the ORd-CiPA model on gfx950.
//
// Replaces, per cell, the reference's SUNDIALS CVODE usage (CVodeCreate(CV_BDF,
// CV_NEWTON) + CVDense + scalar tolerances + CVodeSetMaxStep(dt), output in CV_NORMAL mode
// on the dt grid: CEP7Cell/modules/drugs_herg.cpp:151-158,221-229;
// CEPCell_NEO/modules/cipa_drug_bench.cpp:158-166,213).  The integrator follows CVODE's
// published algorithm (Nordsieck array, fixed-leading-coefficient BDF 1..5, modified
// Newton, DQ Jacobian with CVODE's reuse rules, local error test, order/step selection,
// initial step estimate) -- the same steps as the FP64 CPU oracle.
//
// GPU organisation (one cell per lane):
//   * The per-lane algorithm is a resumable phase machine that stops every time it needs
//     a right-hand side.  The wavefront loop evaluates exactly one RHS per lane per
//     iteration -- a Newton iterate, a Jacobian column or an output point -- so the
//     dominant cost never diverges even though lanes are in different phases (a lane
//     building its 48-column Jacobian does not stall the others).
//   * Newton systems use the Jacobian's structure: 26 Hodgkin-Huxley gates couple only
//     to themselves and V, nca to jca/cass, the 9-state Markov IKr block to V, the three
//     charge accumulators to nothing downstream, D is constant.  Eliminating gates and
//     the Markov block leaves a 12x12 core (CaMKt, 8 concentrations, 2 Jrel, V last).
//     With V ordered last the Markov block (an M-matrix) and the core factor without
//     pivoting; the solution is that of the dense system built from the same
//     finite-difference entries (entries outside the pattern are exactly zero).
//   * Nordsieck history, error weights, the FD Jacobian and the factors live in HBM as
//     SoA (element k of cell c at k*pitch + c); scalars stay in registers.
#pragma once
#include "biomarker.hpp"
#include "ord_model.hpp"

namespace cep {
namespace ordbdf {

constexpr int N = 52;
constexpr int QMAX = 5;
constexpr int NJC = 48;     // Jacobian columns evaluated (all states but D and the accumulators)
// factor storage layout (doubles per cell)
constexpr int NG = 26;      // simple gates
constexpr int F_RG = 0, F_CG = NG, F_NCA = 2 * NG, F_MLU = 2 * NG + 4, F_W = F_MLU + 81, F_SLU = F_W + 9;
constexpr int NFAC = F_SLU + 144;
// scalar save area
constexpr int SD_N = 40;    // doubles
constexpr int SI_N = 24;    // ints

// controller constants (CVODE defaults)
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1;
constexpr double ETACF = 0.25, ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5;
constexpr int SMALL_NST = 10, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10, MAXNEF = 7, MAXNCF = 10;
constexpr int MAXCOR = 3, MSBP = 20, MSBJ = 50;
constexpr double NLSCOEF = 0.1, CRDOWN = 0.3, RDIV = 2.0, DGMAX = 0.3, DGMAX_J = 0.2;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr double UROUND = 2.220446049250313e-16;

enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };
// phases: what the pending RHS request is for
enum { PH_FRESH = 0, PH_INIT_F, PH_HIN, PH_STEP_F0, PH_JAC, PH_NEWTON, PH_OUTPUT, PH_RESTART_F, PH_READY, PH_DONE };

__device__ constexpr int kSimpleGates[NG] = { 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                                              23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 44, 45, 46 };
__device__ constexpr int kCore[12] = { 1, 2, 3, 4, 5, 6, 7, 8, 9, 47, 48, 0 };   // V last
constexpr int CORE_CASS = 1, CORE_V = 11;

CEP_HD constexpr int jslot(int s) { return s < 43 ? s : s - 1; }
CEP_HD constexpr int jstate(int slot) { return slot < 43 ? slot : slot + 1; }

// which gate states each core row couples to (currents: INa 10-15, INaL 16-18, Ito 19-24,
// ICaL/ICaNa/ICaK 25-33, IKs 44-45, IK1 46; IKr through the Markov O state 38)
CEP_HD constexpr bool couples(int row, int g)
{
  const bool na = g >= 10 && g <= 15, nal = g >= 16 && g <= 18, to = g >= 19 && g <= 24;
  const bool cal = g >= 25 && g <= 33, ks = g == 44 || g == 45, k1 = g == 46, o = g == 38;
  switch (row) {
    case 0: return na || nal || to || cal || ks || k1 || o;
    case 3: return na || nal;
    case 5: return to || ks || k1 || o;
    case 2: case 4: case 6: case 47: case 48: return cal;
    case 49: return nal || to || cal || ks || k1 || o;
    case 50: return nal;
    case 51: return cal;
    default: return false;
  }
}

struct Buffers {
  int ncell, pitch;
  double *y;        // [52] current output state (CellML STATES)
  double *zn;       // [6*52]
  double *ewt, *acor, *f0;   // [52]
  double *jd;       // [48*52] FD Jacobian, column slot-major
  double *fac;      // [NFAC]
  double *sd;       // [SD_N] scalar save
  int *si;          // [SI_N]
  double *prm;      // [ord::NP] per-cell parameters
  double *bmd;      // [BM_TOTAL_D]
  int *bmi;         // [BM_I_N]
  double *ct, *cv;  // Ca trace [cap]
  double *vtrace;   // optional Vm trace of the last beat [tcap], or null
  int tcap;
  long long *stats; // [8] nst nfe nje nsetups nni ncfn netf nout
  int *status;
  double dt, rtol, atol, uround_fd;
  int maxord, nwin_unused;
  int cap;
};

struct Param {
  const double *p;
  int pitch;
  CEP_D double operator()(int k) const { return p[(size_t)k * pitch]; }
};

template <typename R>
struct Machine {
  const Buffers &B;
  const int c;
  const Param P;
  // scalars
  double tn, h, hprime, eta, hscale, etamax, rl1, gamma, gammap, gamrat, crate, acnrm, del, delp;
  double saved_t, tq1, tq2, tq3, tq4, tq5, saved_tq5, tout, hg, hub, hlb, hnew, hs;
  double tau[7], l[6];
  int q, qprime, qwait, jcur, phase, jcol, callSetup, convfail, nflag, ncf, nef, mnewt;
  int iout, iout_end, status, hin_count, hin_ok, nstloc;
  long long nst, nstlp, nstlj, nfe, nje, nsetups, nni, ncfn, netf;
  double inc_cur;

  CEP_D Machine(const Buffers &b, int cell) : B(b), c(cell), P{b.prm + cell, b.pitch} {}

  // SoA accessors
  CEP_D double &ZN(int j, int i) const { return B.zn[(size_t)(j * N + i) * B.pitch + c]; }
  CEP_D double &EWT(int i) const { return B.ewt[(size_t)i * B.pitch + c]; }
  CEP_D double &ACOR(int i) const { return B.acor[(size_t)i * B.pitch + c]; }
  CEP_D double &F0(int i) const { return B.f0[(size_t)i * B.pitch + c]; }
  CEP_D double &JD(int slot, int row) const { return B.jd[(size_t)(slot * N + row) * B.pitch + c]; }
  CEP_D double J(int row, int col) const { return JD(jslot(col), row); }
  CEP_D double &FAC(int k) const { return B.fac[(size_t)k * B.pitch + c]; }
  CEP_D double &Y(int i) const { return B.y[(size_t)i * B.pitch + c]; }

  // ---------------- persistence
  CEP_D void load()
  {
    const double *d = B.sd + c;
    const size_t p = B.pitch;
    double *dv[SD_N] = { &tn, &h, &hprime, &eta, &hscale, &etamax, &rl1, &gamma, &gammap, &gamrat,
                         &crate, &acnrm, &del, &delp, &saved_t, &tq1, &tq2, &tq3, &tq4, &tq5,
                         &saved_tq5, &tout, &hg, &hub, &hlb, &hnew, &hs, &tau[0], &tau[1], &tau[2],
                         &tau[3], &tau[4], &tau[5], &tau[6], &l[0], &l[1], &l[2], &l[3], &l[4], &l[5] };
#pragma unroll
    for (int k = 0; k < SD_N; ++k) *dv[k] = d[k * p];
    const int *s = B.si + c;
    int *iv[15] = { &q, &qprime, &qwait, &jcur, &phase, &jcol, &callSetup, &convfail, &nflag, &ncf,
                    &nef, &mnewt, &iout, &hin_count, &hin_ok };
#pragma unroll
    for (int k = 0; k < 15; ++k) *iv[k] = s[k * p];
    nst = s[15 * p]; nstlp = s[16 * p]; nstlj = s[17 * p];
    status = B.status[c];
    const long long *st = B.stats + c;
    nfe = st[1 * p]; nje = st[2 * p]; nsetups = st[3 * p]; nni = st[4 * p]; ncfn = st[5 * p]; netf = st[6 * p];
  }
  CEP_D void store() const
  {
    double *d = B.sd + c;
    const size_t p = B.pitch;
    const double dv[SD_N] = { tn, h, hprime, eta, hscale, etamax, rl1, gamma, gammap, gamrat, crate,
                              acnrm, del, delp, saved_t, tq1, tq2, tq3, tq4, tq5, saved_tq5, tout,
                              hg, hub, hlb, hnew, hs, tau[0], tau[1], tau[2], tau[3], tau[4], tau[5],
                              tau[6], l[0], l[1], l[2], l[3], l[4], l[5] };
#pragma unroll
    for (int k = 0; k < SD_N; ++k) d[k * p] = dv[k];
    int *s = B.si + c;
    const int iv[15] = { q, qprime, qwait, jcur, phase, jcol, callSetup, convfail, nflag, ncf, nef,
                         mnewt, iout, hin_count, hin_ok };
#pragma unroll
    for (int k = 0; k < 15; ++k) s[k * p] = iv[k];
    s[15 * p] = (int)nst; s[16 * p] = (int)nstlp; s[17 * p] = (int)nstlj;
    B.status[c] = status;
    long long *st = B.stats + c;
    st[0] = nst; st[1 * p] = nfe; st[2 * p] = nje; st[3 * p] = nsetups; st[4 * p] = nni;
    st[5 * p] = ncfn; st[6 * p] = netf; st[7 * p] = iout;
  }

  // ---------------- small helpers
  CEP_D double wrms_reg(const double *x) const
  {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) { const double t = x[i] * EWT(i); s += t * t; }
    return sqrt(s * (1.0 / N));
  }
  CEP_D void set_ewt_from_zn0()
  {
#pragma unroll
    for (int i = 0; i < N; ++i) EWT(i) = 1.0 / (B.rtol * fabs(ZN(0, i)) + B.atol);
  }

  CEP_D void predict()
  {
    tn += h;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double z[QMAX + 1];
#pragma unroll
      for (int j = 0; j <= QMAX; ++j) z[j] = (j <= q) ? ZN(j, i) : 0.0;
#pragma unroll
      for (int k = 1; k <= QMAX; ++k)
#pragma unroll
        for (int j = QMAX; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] += z[j];
#pragma unroll
      for (int j = 0; j <= QMAX; ++j) if (j < q) ZN(j, i) = z[j];
    }
  }
  __device__ __noinline__ void restore()
  {
    tn = saved_t;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double z[QMAX + 1];
#pragma unroll
      for (int j = 0; j <= QMAX; ++j) z[j] = (j <= q) ? ZN(j, i) : 0.0;
#pragma unroll
      for (int k = 1; k <= QMAX; ++k)
#pragma unroll
        for (int j = QMAX; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] -= z[j];
#pragma unroll
      for (int j = 0; j <= QMAX; ++j) if (j < q) ZN(j, i) = z[j];
    }
  }
  __device__ __noinline__ void rescale()
  {
    double fac[QMAX + 1];
    fac[0] = 1.0;
#pragma unroll
    for (int j = 1; j <= QMAX; ++j) fac[j] = fac[j - 1] * eta;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 1; j <= QMAX; ++j)
        if (j <= q) ZN(j, i) *= fac[j];
    h = hscale * eta;
    hscale = h;
  }

  CEP_D void set_bdf()
  {
    double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0, hsum = h;
    l[0] = l[1] = 1.0;
#pragma unroll
    for (int i = 2; i <= QMAX; ++i) l[i] = 0.0;
    if (q > 1) {
      for (int j = 2; j < q; ++j) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        alpha0 -= 1.0 / j;
#pragma unroll
        for (int i = QMAX; i >= 1; --i) if (i <= j) l[i] += l[i - 1] * xi_inv;
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = QMAX; i >= 1; --i) if (i <= q) l[i] += l[i - 1] * xistar_inv;
    }
    // error-test / order-change coefficients
    const double lq = l[q < 1 ? 1 : (q > QMAX ? QMAX : q)];
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq2 = fabs(A1 / (alpha0 * A2));
    tq5 = fabs(A2 * xistar_inv / (lq * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double Cc = xistar_inv / lq;
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq1 = fabs(Cc * Cpinv);
      } else {
        tq1 = 1.0;
      }
      const double hsum2 = hsum + tau[q];
      const double xi2 = h / hsum2;
      const double A5 = alpha0 - (1.0 / (q + 1));
      const double A6 = alpha0_hat - xi2;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq3 = fabs(Cppinv / (xi2 * (q + 2) * A5));
    }
    tq4 = NLSCOEF / tq2;
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  CEP_D void increase_order()
  {
    double ll[QMAX + 1];
#pragma unroll
    for (int i = 0; i <= QMAX; ++i) ll[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = hscale;
    ll[2] = 1.0;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
#pragma unroll
        for (int i = QMAX; i >= 2; --i) if (i <= j + 2) ll[i] = ll[i] * xiold + ll[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
    // zn[q+1] = A1 * zn[qmax] (acor saved there); zn[j] += l[j] * zn[q+1]
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double zl = A1 * ZN(QMAX, i);
#pragma unroll
      for (int j = 2; j <= QMAX; ++j) if (j <= q) ZN(j, i) += ll[j] * zl;
#pragma unroll
      for (int j = 2; j <= QMAX; ++j) if (j == q + 1) ZN(j, i) = zl;
    }
  }
  CEP_D void decrease_order()
  {
    double ll[QMAX + 1];
#pragma unroll
    for (int i = 0; i <= QMAX; ++i) ll[i] = 0.0;
    ll[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      const double xi = hsum / hscale;
#pragma unroll
      for (int i = QMAX; i >= 2; --i) if (i <= j + 2) ll[i] = ll[i] * xi + ll[i - 1];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double zq = 0.0;
#pragma unroll
      for (int j = 0; j <= QMAX; ++j) if (j == q) zq = ZN(j, i);
#pragma unroll
      for (int j = 2; j < QMAX; ++j) if (j < q) ZN(j, i) -= ll[j] * zq;
    }
  }
  CEP_D void adjust_order(int dq)
  {
    if ((q == 2) && (dq != 1)) return;
    if (dq == 1) increase_order(); else if (dq == -1) decrease_order();
  }

  // ---------------- linear algebra: structured factorisation of M = I - gamma*J
  CEP_D bool factor()
  {
    const double g = gamma;
    // simple gates
#pragma unroll
    for (int k = 0; k < NG; ++k) {
      const int s = kSimpleGates[k];
      const double r = 1.0 / (1.0 - g * J(s, s));
      FAC(F_RG + k) = r;
      FAC(F_CG + k) = (-g * J(s, 0)) * r;
    }
    // nca: row {cass, jca, nca}
    const double rn = 1.0 / (1.0 - g * J(33, 33));
    const double e_nj = rn * (-g * J(33, 30));
    const double c_jca = FAC(F_CG + 20);   // jca is simple gate #20
    FAC(F_NCA + 0) = rn;
    FAC(F_NCA + 1) = e_nj;
    FAC(F_NCA + 2) = -e_nj * c_jca;            // cnV
    FAC(F_NCA + 3) = rn * (-g * J(33, 2));     // cnC
    // Markov 9x9 LU (no pivoting: column diagonally dominant M-matrix)
    double A[9][9];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = 0; j < 9; ++j) A[i][j] = (i == j ? 1.0 : 0.0) - g * J(34 + i, 34 + j);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double pinv = 1.0 / A[k][k];
#pragma unroll
      for (int i = k + 1; i < 9; ++i) {
        A[i][k] *= pinv;
#pragma unroll
        for (int j = k + 1; j < 9; ++j) A[i][j] -= A[i][k] * A[k][j];
      }
    }
    double w[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) w[i] = -g * J(34 + i, 0);
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j) w[i] -= A[i][j] * w[j];
#pragma unroll
    for (int i = 8; i >= 0; --i) {
#pragma unroll
      for (int j = i + 1; j < 9; ++j) w[i] -= A[i][j] * w[j];
      w[i] /= A[i][i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
#pragma unroll
      for (int j = 0; j < 9; ++j) FAC(F_MLU + i * 9 + j) = A[i][j];
      FAC(F_W + i) = w[i];
    }
    // core Schur complement (12x12, V last)
    double S[12][12];
#pragma unroll
    for (int a = 0; a < 12; ++a)
#pragma unroll
      for (int b = 0; b < 12; ++b) S[a][b] = (a == b ? 1.0 : 0.0) - g * J(kCore[a], kCore[b]);
    const double cnV = FAC(F_NCA + 2), cnC = FAC(F_NCA + 3);
#pragma unroll
    for (int a = 0; a < 12; ++a) {
      const int row = kCore[a];
      double sv = 0.0;
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(row, kSimpleGates[k])) sv += (-g * J(row, kSimpleGates[k])) * FAC(F_CG + k);
      if (couples(row, 33)) {
        const double mn = -g * J(row, 33);
        sv += mn * cnV;
        S[a][CORE_CASS] -= mn * cnC;
      }
      if (couples(row, 38)) sv += (-g * J(row, 38)) * w[4];
      S[a][CORE_V] -= sv;
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double pinv = 1.0 / S[k][k];
#pragma unroll
      for (int i = k + 1; i < 12; ++i) {
        S[i][k] *= pinv;
#pragma unroll
        for (int j = k + 1; j < 12; ++j) S[i][j] -= S[i][k] * S[k][j];
      }
    }
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 12; ++a) {
      ok = ok && (S[a][a] != 0.0) && isfinite(S[a][a]);
#pragma unroll
      for (int b = 0; b < 12; ++b) FAC(F_SLU + a * 12 + b) = S[a][b];
    }
    return ok;
  }

  // solve M x = b in place (b in registers), then CVODE's 2/(1+gamrat) scaling
  CEP_D void solve(double *b) const
  {
    const double g = gammap;
    double beta[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) beta[k] = FAC(F_RG + k) * b[kSimpleGates[k]];
    const double beta_n = FAC(F_NCA + 0) * b[33] - FAC(F_NCA + 1) * beta[20];
    double bm[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) bm[i] = b[34 + i];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j) bm[i] -= FAC(F_MLU + i * 9 + j) * bm[j];
#pragma unroll
    for (int i = 8; i >= 0; --i) {
#pragma unroll
      for (int j = i + 1; j < 9; ++j) bm[i] -= FAC(F_MLU + i * 9 + j) * bm[j];
      bm[i] /= FAC(F_MLU + i * 9 + i);
    }
    double bk[12];
#pragma unroll
    for (int a = 0; a < 12; ++a) {
      const int row = kCore[a];
      double s = b[row];
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(row, kSimpleGates[k])) s += g * J(row, kSimpleGates[k]) * beta[k];
      if (couples(row, 33)) s += g * J(row, 33) * beta_n;
      if (couples(row, 38)) s += g * J(row, 38) * bm[4];
      bk[a] = s;
    }
#pragma unroll
    for (int i = 0; i < 12; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j) bk[i] -= FAC(F_SLU + i * 12 + j) * bk[j];
#pragma unroll
    for (int i = 11; i >= 0; --i) {
#pragma unroll
      for (int j = i + 1; j < 12; ++j) bk[i] -= FAC(F_SLU + i * 12 + j) * bk[j];
      bk[i] /= FAC(F_SLU + i * 12 + i);
    }
    const double xV = bk[CORE_V], xC = bk[CORE_CASS];
#pragma unroll
    for (int a = 0; a < 12; ++a) b[kCore[a]] = bk[a];
#pragma unroll
    for (int k = 0; k < NG; ++k) b[kSimpleGates[k]] = beta[k] - FAC(F_CG + k) * xV;
    b[33] = beta_n - FAC(F_NCA + 2) * xV - FAC(F_NCA + 3) * xC;
#pragma unroll
    for (int i = 0; i < 9; ++i) b[34 + i] = bm[i] - FAC(F_W + i) * xV;
    // accumulators: x = b + gamma * J(acc, :) x
#pragma unroll
    for (int r = 49; r <= 51; ++r) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 12; ++a) s += J(r, kCore[a]) * b[kCore[a]];
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(r, kSimpleGates[k])) s += J(r, kSimpleGates[k]) * b[kSimpleGates[k]];
      if (couples(r, 33)) s += J(r, 33) * b[33];
      if (couples(r, 38)) s += J(r, 38) * b[38];
      b[r] += g * s;
    }
    if (gamrat != 1.0) {
      const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
      for (int i = 0; i < N; ++i) b[i] *= sc;
    }
  }

  // ---------------- phase machine.  Actions chain inside one dispatcher loop (drive) so
  // every phase body is instantiated once; ACT_REQ means (treq, yreq) holds the next RHS
  // request, ACT_DONE that the lane has finished this launch.
  enum { ACT_DONE = 0, ACT_REQ, ACT_OUTPUT_CHECK, ACT_STEP_ATTEMPT, ACT_SETUP, ACT_NEWTON_START,
         ACT_NEWTON_ITER, ACT_ERROR_TEST, ACT_CONV_FAIL, ACT_ERR_FAIL, ACT_HIN_FINISH, ACT_REQ_F0 };

  // tout reached: interpolate y(tout) and request an output RHS; else start a new step
  CEP_D int output_check(double *yreq, double &treq)
  {
    if ((tn - tout) * h >= 0.0) {
      const double s = (tout - tn) / h;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = QMAX; j >= 0; --j) if (j <= q) acc = ZN(j, i) + s * acc;
        yreq[i] = acc;
      }
      treq = tout;
      phase = PH_OUTPUT;
      return ACT_REQ;
    }
    if (nst > 0) set_ewt_from_zn0();
    if (++nstloc > 500) { status = ST_MAX_STEPS; phase = PH_DONE; return ACT_DONE; }
    saved_t = tn;
    ncf = nef = 0;
    nflag = FIRST_CALL;
    if ((nst > 0) && (hprime != h)) {
      if (qprime != q) {
        adjust_order(qprime - q);
        q = qprime; qwait = q + 1;
      }
      rescale();
    }
    return ACT_STEP_ATTEMPT;
  }

  CEP_D int step_attempt()
  {
    predict();
    set_bdf();
    convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? NO_FAILURES : FAIL_OTHER;
    callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) ||
                (nst >= nstlp + MSBP) || (fabs(gamrat - 1.0) > DGMAX);
    return ACT_REQ_F0;
  }

  CEP_D int request_f0(double *yreq, double &treq)
  {
#pragma unroll
    for (int i = 0; i < N; ++i) yreq[i] = ZN(0, i);
    treq = tn;
    phase = PH_STEP_F0;
    return ACT_REQ;
  }

  CEP_D void jac_request(double *yreq, double &treq)
  {
    const int s = jstate(jcol);
    double ys = 0.0, ew = 1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double z = ZN(0, i);
      yreq[i] = z;
      if (i == s) { ys = z; ew = EWT(i); }
    }
    const double srur = sqrt(B.uround_fd);
    double inc = srur * fabs(ys);
    const double alt = hg / ew;   // hg holds CVODE's minInc during the Jacobian phase
    if (alt > inc) inc = alt;
    inc_cur = inc;
#pragma unroll
    for (int i = 0; i < N; ++i) if (i == s) yreq[i] = ys + inc;
    treq = tn;
    phase = PH_JAC;
  }

  // consume f = f(tn, zn0) of a step attempt
  CEP_D int on_step_f0(const double *f, double *yreq, double &treq)
  {
    nfe++;
    if (!callSetup) return ACT_NEWTON_START;
    const double dgamma = fabs((gamma / gammap) - 1.0);
    const bool jbad = (nst == 0) || (nst > nstlj + MSBJ) ||
                      ((convfail == FAIL_BAD_J) && (dgamma < DGMAX_J)) || (convfail == FAIL_OTHER);
    if (!jbad) { jcur = 0; return ACT_SETUP; }
#pragma unroll
    for (int i = 0; i < N; ++i) F0(i) = f[i];
    nstlj = nst;
    jcur = 1;
    const double fnorm = wrms_reg(f);
    hg = (fnorm != 0.0) ? (1000.0 * fabs(h) * B.uround_fd * N * fnorm) : 1.0;
    jcol = 0;
    jac_request(yreq, treq);
    return ACT_REQ;
  }

  // consume one Jacobian column; after the last, f is reloaded with f(tn, zn0)
  CEP_D int on_jac(double *f, double *yreq, double &treq)
  {
    nfe++;
    const double iinv = 1.0 / inc_cur;
#pragma unroll
    for (int i = 0; i < N; ++i) JD(jcol, i) = iinv * f[i] - iinv * F0(i);
    jcol++;
    if (jcol < NJC) { jac_request(yreq, treq); return ACT_REQ; }
    nje++;
#pragma unroll
    for (int i = 0; i < N; ++i) f[i] = F0(i);
    return ACT_SETUP;
  }

  CEP_D int setup()
  {
    const bool ok = factor();
    nsetups++;
    callSetup = 0;
    gamrat = crate = 1.0;
    gammap = gamma;
    nstlp = nst;
    return ok ? ACT_NEWTON_START : ACT_CONV_FAIL;
  }

  CEP_D int newton_iter(const double *ftemp, double *yreq, double &treq)
  {
    double b[N];
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = gamma * ftemp[i] - (rl1 * ZN(1, i) + ACOR(i));
    solve(b);
    nni++;
    del = wrms_reg(b);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double a = ACOR(i) + b[i];
      ACOR(i) = a;
      yreq[i] = ZN(0, i) + a;
    }
    if (mnewt > 0) { const double r = del / delp; crate = (CRDOWN * crate > r) ? CRDOWN * crate : r; }
    const double dcon = del * (crate < 1.0 ? crate : 1.0) / tq4;
    if (dcon <= 1.0) {
      if (mnewt == 0) {
        acnrm = del;
      } else {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) { const double t = ACOR(i) * EWT(i); s += t * t; }
        acnrm = sqrt(s * (1.0 / N));
      }
      jcur = 0;
      return ACT_ERROR_TEST;
    }
    mnewt++;
    if ((mnewt == MAXCOR) || ((mnewt >= 2) && (del > RDIV * delp))) {
      if (!jcur) {   // retry the nonlinear solve with a fresh Jacobian
        callSetup = 1;
        convfail = FAIL_BAD_J;
        return ACT_REQ_F0;
      }
      return ACT_CONV_FAIL;
    }
    delp = del;
    treq = tn;
    phase = PH_NEWTON;
    return ACT_REQ;
  }

  CEP_D int conv_fail()
  {
    ncfn++;
    restore();
    ncf++;
    etamax = 1.0;
    if (ncf == MAXNCF) { status = ST_CONV_FAIL; phase = PH_DONE; return ACT_DONE; }
    eta = ETACF;
    nflag = PREV_CONV_FAIL;
    rescale();
    return ACT_STEP_ATTEMPT;
  }

  CEP_D int error_test()
  {
    const double dsm = acnrm * tq2;
    if (dsm <= 1.0) {
      complete_step();
      prepare_next(dsm);
      etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
      return ACT_OUTPUT_CHECK;
    }
    nef++;
    netf++;
    nflag = PREV_ERR_FAIL;
    if (nef == MAXNEF || !(dsm == dsm)) { restore(); status = ST_ERR_FAIL; phase = PH_DONE; return ACT_DONE; }
    etamax = 1.0;
    restore();
    if (nef <= MXNEF1) {
      double e = 1.0 / (pow(BIAS2 * dsm, 1.0 / (q + 1)) + ADDON);
      if (e < ETAMIN) e = ETAMIN;
      if (nef >= SMALL_NEF && e > ETAMXF) e = ETAMXF;
      eta = e;
      rescale();
      return ACT_STEP_ATTEMPT;
    }
    if (q > 1) {
      eta = ETAMIN;
      adjust_order(-1);
      q--;
      qwait = q + 1;
      rescale();
      return ACT_STEP_ATTEMPT;
    }
    // order 1: restart from f(tn, zn0) with a smaller step
    eta = ETAMIN;
    h *= eta;
    hscale = h;
    qwait = LONG_WAIT;
    return ACT_ERR_FAIL;
  }

  CEP_D int drive(double *f, double *yreq, double &treq, int act)
  {
    while (act > ACT_REQ) {
      switch (act) {
        case ACT_OUTPUT_CHECK: act = output_check(yreq, treq); break;
        case ACT_STEP_ATTEMPT: act = step_attempt(); break;
        case ACT_REQ_F0: act = request_f0(yreq, treq); break;
        case ACT_SETUP: act = setup(); break;
        case ACT_NEWTON_START:
#pragma unroll
          for (int i = 0; i < N; ++i) ACOR(i) = 0.0;
          mnewt = 0;
          del = delp = 0.0;
          act = ACT_NEWTON_ITER;
          break;
        case ACT_NEWTON_ITER: act = newton_iter(f, yreq, treq); break;
        case ACT_ERROR_TEST: act = error_test(); break;
        case ACT_CONV_FAIL: act = conv_fail(); break;
        case ACT_ERR_FAIL:
#pragma unroll
          for (int i = 0; i < N; ++i) yreq[i] = ZN(0, i);
          treq = tn;
          phase = PH_RESTART_F;
          act = ACT_REQ;
          break;
        case ACT_HIN_FINISH: act = hin_finish(); break;
        default: act = ACT_DONE; break;
      }
    }
    return act;
  }

  // consume the RHS just evaluated for the pending phase and run to the next request
  CEP_D bool consume(double *f, double *yreq, double &treq)
  {
    int act = ACT_DONE;
    switch (phase) {
      case PH_INIT_F:
#pragma unroll
        for (int i = 0; i < N; ++i) ZN(1, i) = f[i];
        nfe = 1;
        set_ewt_from_zn0();
        act = hin_begin(yreq, treq);
        break;
      case PH_HIN: act = on_hin(f, yreq, treq); break;
      case PH_STEP_F0: act = on_step_f0(f, yreq, treq); break;
      case PH_JAC: act = on_jac(f, yreq, treq); break;
      case PH_NEWTON: nfe++; act = ACT_NEWTON_ITER; break;
      case PH_RESTART_F:
        nfe++;
#pragma unroll
        for (int i = 0; i < N; ++i) ZN(1, i) = h * f[i];
        act = ACT_STEP_ATTEMPT;
        break;
      case PH_READY: act = ACT_OUTPUT_CHECK; break;
      default: act = ACT_DONE; break;
    }
    return drive(f, yreq, treq, act) == ACT_REQ;
  }

  CEP_D void complete_step()
  {
    nst++;
#pragma unroll
    for (int i = QMAX; i >= 2; --i) if (i <= q) tau[i] = tau[i - 1];
    if ((q == 1) && (nst > 1)) tau[2] = tau[1];
    tau[1] = h;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double a = ACOR(i);
#pragma unroll
      for (int j = 0; j <= QMAX; ++j) if (j <= q) ZN(j, i) += l[j] * a;
    }
    qwait--;
    if ((qwait == 1) && (q != B.maxord)) {
#pragma unroll
      for (int i = 0; i < N; ++i) ZN(QMAX, i) = ACOR(i);
      saved_tq5 = tq5;
    }
  }

  CEP_D void set_eta()
  {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      if (eta > etamax) eta = etamax;
      const double r = fabs(h) * (1.0 / B.dt) * eta;
      eta /= (r > 1.0 ? r : 1.0);
      hprime = h * eta;
    }
  }

  CEP_D void prepare_next(double dsm)
  {
    if (etamax == 1.0) {
      if (qwait < 2) qwait = 2;
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    const double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / (q + 1)) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    double etaqm1 = 0.0, etaqp1 = 0.0;
    if (q > 1) {
      double zq[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j <= QMAX; ++j) if (j == q) v = ZN(j, i);
        zq[i] = v;
      }
      const double ddn = wrms_reg(zq) * tq1;
      etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    if (q != B.maxord && saved_tq5 != 0.0) {
      const double cquot = (tq5 / saved_tq5) * pow(h / tau[2], (double)(q + 1));
      double tv[N];
#pragma unroll
      for (int i = 0; i < N; ++i) tv[i] = ACOR(i) - cquot * ZN(QMAX, i);
      const double dup = wrms_reg(tv) * tq3;
      etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (q + 2)) + ADDON);
    }
    double etam = etaq;
    if (etaqm1 > etam) etam = etaqm1;
    if (etaqp1 > etam) etam = etaqp1;
    if (etam < THRESH) {
      eta = 1.0; qprime = q;
    } else if (etam == etaq) {
      eta = etaq; qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1; qprime = q - 1;
    } else {
      eta = etaqp1; qprime = q + 1;
#pragma unroll
      for (int i = 0; i < N; ++i) ZN(QMAX, i) = ACOR(i);
    }
    set_eta();
  }

  // ---- initial step size (CVODE's cvHin), one RHS per trial step
  CEP_D int hin_begin(double *yreq, double &treq)
  {
    const double tdist = fabs(B.dt);     // first output point is t0 + dt
    const double tround = UROUND * fabs(tout);
    hlb = HLB_FACTOR * tround;
    double hub_inv = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double tol = HUB_FACTOR * fabs(ZN(0, i)) + B.atol;
      const double r = fabs(ZN(1, i)) / tol;
      if (r > hub_inv) hub_inv = r;
    }
    hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    hg = sqrt(hlb * hub);
    if (hub < hlb) { hnew = hg; hin_ok = 2; return ACT_HIN_FINISH; }
    hin_count = 1;
    hin_ok = 0;
    hs = hg;
    return hin_request(yreq, treq);
  }
  CEP_D int hin_request(double *yreq, double &treq)
  {
#pragma unroll
    for (int i = 0; i < N; ++i) yreq[i] = hg * ZN(1, i) + ZN(0, i);
    treq = tn + hg;
    phase = PH_HIN;
    return ACT_REQ;
  }
  CEP_D int on_hin(const double *f, double *yreq, double &treq)
  {
    nfe++;
    double d[N];
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = (f[i] - ZN(1, i)) * (1.0 / hg);
    const double yddnrm = wrms_reg(d);
    hs = hg;
    if (hin_ok || hin_count == 4) { hnew = hg; hin_ok = 1; return ACT_HIN_FINISH; }
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    const double hrat = hnew / hg;
    if ((hrat > 0.5) && (hrat < 2.0)) hin_ok = 1;
    if ((hin_count > 1) && (hrat > 2.0)) { hnew = hg; hin_ok = 1; }
    hg = hnew;
    hin_count++;
    return hin_request(yreq, treq);
  }
  // hin_ok == 2: hub < hlb, take hg directly
  CEP_D int hin_finish()
  {
    const bool direct = hin_ok == 2;
    double h0 = direct ? hnew : H_BIAS * hnew;
    if (!direct) {
      if (h0 < hlb) h0 = hlb;
      if (h0 > hub) h0 = hub;
    }
    const double rh = fabs(h0) / B.dt;   // hmax = dt
    if (rh > 1.0) h0 /= rh;
    h = h0; hscale = h0; hprime = h0;
#pragma unroll
    for (int i = 0; i < N; ++i) ZN(1, i) *= h0;
    return ACT_OUTPUT_CHECK;
  }
};

}  // namespace ordbdf
}  // namespace cep
