/*
 * adams.c -- ORACLE (test infrastructure only): the variable-order, variable-step Adams
 * PECE integrator of Shampine & Gordon ("Computer Solution of Ordinary Differential
 * Equations", Freeman 1975: DE / STEP / INTRP), as the reference carries it in
 *   libs/ode.cpp  (ode :589-805, de :14-403, step :970-1645, intrp :446-586)
 * and uses it in CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:119-146.
 *
 * One adams_mem_t holds what the reference keeps in WORK(100+21*NEQN) / IWORK(5): the modified
 * divided differences phi (16 columns: orders 1..12, the two next differences and two
 * rounding-compensation columns), the step-size history psi, the coefficients
 * alpha / beta / sigma / v / w / g, and the scalars x, h, hold, told, delsgn, ns, k, kold,
 * isnold, start, phase1, nornd.  Every arithmetic step follows the published algorithm in
 * its order, so the results agree with the reference's compiled ode.cpp (pinned by
 * tests/test_adams.py against vectors from oracle/_ref/ref_ode).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ADAMS_COLS 16
#define ADAMS_MAXNUM 500
static const double kUround = 2.220446049250313E-016;
/* error-constant and power-of-two tables of STEP */
static const double kGstr[13] = { 0.50, 0.0833, 0.0417, 0.0264, 0.0188, 0.0143, 0.0114, 0.00936,
                                  0.00789, 0.00679, 0.00592, 0.00524, 0.00468 };
static const double kTwo[13] = { 2.0, 4.0, 8.0, 16.0, 32.0, 64.0, 128.0, 256.0, 512.0, 1024.0,
                                 2048.0, 4096.0, 8192.0 };

struct adams_mem {
  int n;
  adams_rhs_fn f;
  void *user;
  double *phi, *yy, *wt, *p, *yp, *ypout;
  double psi[12], alpha[12], beta[12], sig[13], v[12], w[12], g[13];
  double x, h, hold, told, delsgn;
  int ns, k, kold, isnold;
  int start, phase1, nornd;
  long nsteps, nfe;
};

static double sgn(double a) { return a < 0.0 ? -1.0 : 1.0; }
static double dmax(double a, double b) { return b < a ? a : b; }
static double dmin(double a, double b) { return b < a ? b : a; }

adams_mem_t *adams_create(int n, adams_rhs_fn f, void *user)
{
  adams_mem_t *m = (adams_mem_t *)calloc(1, sizeof(adams_mem_t));
  m->n = n;
  m->f = f;
  m->user = user;
  m->phi = (double *)calloc((size_t)n * ADAMS_COLS, sizeof(double));
  m->yy = (double *)calloc((size_t)n * 5, sizeof(double));
  m->wt = m->yy + n;
  m->p = m->wt + n;
  m->yp = m->p + n;
  m->ypout = m->yp + n;
  return m;
}

void adams_free(adams_mem_t *m)
{
  if (!m) return;
  free(m->phi);
  free(m->yy);
  free(m);
}

void adams_counts(const adams_mem_t *m, long *nsteps, long *nfe, int *order)
{
  if (nsteps) *nsteps = m->nsteps;
  if (nfe) *nfe = m->nfe;
  if (order) *order = m->kold;
}

#define PHI(i, col) m->phi[(i) + (size_t)(col) * n]   /* col 0-based: column col+1 of the method */

static void rhs(adams_mem_t *m, double t, const double *y, double *yp)
{
  m->f(t, y, yp, m->user);
  m->nfe++;
}

/* INTRP: the interpolating polynomial at xout from the differences of the last step */
static void interpolate(adams_mem_t *m, double xout, double *yout)
{
  const int n = m->n;
  double g[13], rho[13], w[13];
  const double hi = xout - m->x;
  const int ki = m->kold + 1;
  for (int i = 1; i <= ki; ++i) w[i - 1] = 1.0 / (double)i;
  g[0] = 1.0;
  rho[0] = 1.0;
  double term = 0.0;
  for (int j = 2; j <= ki; ++j) {
    const double psijm1 = m->psi[j - 2];
    const double gamma = (hi + term) / psijm1;
    const double eta = hi / psijm1;
    for (int i = 1; i <= ki + 1 - j; ++i) w[i - 1] = gamma * w[i - 1] - eta * w[i];
    g[j - 1] = w[0];
    rho[j - 1] = gamma * rho[j - 2];
    term = psijm1;
  }
  for (int l = 0; l < n; ++l) { m->ypout[l] = 0.0; yout[l] = 0.0; }
  for (int j = 1; j <= ki; ++j) {
    const int i = ki + 1 - j;
    for (int l = 0; l < n; ++l) {
      yout[l] = yout[l] + g[i - 1] * PHI(l, i - 1);
      m->ypout[l] = m->ypout[l] + rho[i - 1] * PHI(l, i - 1);
    }
  }
  for (int l = 0; l < n; ++l) yout[l] = m->yy[l] + hi * yout[l];
}

/* STEP: one Adams PECE step of order k with local extrapolation, step-size and order
   control; returns crash = 1 when h or eps must grow (the caller reports iflag 3) */
static int adams_step(adams_mem_t *m, double *eps)
{
  const int n = m->n;
  double *y = m->yy, *wt = m->wt, *p = m->p, *yp = m->yp;
  const double twou = 2.0 * kUround, fouru = 2.0 * twou;
  double erk, erkm1, erkm2, err, absh, temp1, temp2;
  int knew, kp1, kp2, km1, km2, ifail;

  if (fabs(m->h) < fouru * fabs(m->x)) {
    m->h = fouru * fabs(m->x) * sgn(m->h);
    return 1;
  }
  const double p5eps = 0.5 * *eps;
  double round = 0.0;
  for (int l = 0; l < n; ++l) round = round + pow(y[l] / wt[l], 2);
  round = twou * sqrt(round);
  if (p5eps < round) {
    *eps = 2.0 * round * (1.0 + fouru);
    return 1;
  }
  m->g[0] = 1.0;
  m->g[1] = 0.5;
  m->sig[0] = 1.0;
  if (m->start) {
    /* first step: initial step size from the derivative, order 1 */
    rhs(m, m->x, y, yp);
    for (int l = 0; l < n; ++l) { PHI(l, 0) = yp[l]; PHI(l, 1) = 0.0; }
    double total = 0.0;
    for (int l = 0; l < n; ++l) total = total + pow(yp[l] / wt[l], 2);
    total = sqrt(total);
    absh = fabs(m->h);
    if (*eps < 16.0 * total * m->h * m->h) absh = 0.25 * sqrt(*eps / total);
    m->h = dmax(absh, fouru * fabs(m->x)) * sgn(m->h);
    m->hold = 0.0;
    m->k = 1;
    m->kold = 0;
    m->start = 0;
    m->phase1 = 1;
    m->nornd = 1;
    if (p5eps <= 100.0 * round) {
      m->nornd = 0;
      for (int l = 0; l < n; ++l) PHI(l, 14) = 0.0;
    }
  }
  ifail = 0;
  for (;;) {
    const double h = m->h;
    const int k = m->k;
    kp1 = k + 1; kp2 = k + 2; km1 = k - 1; km2 = k - 2;
    /* coefficients for the current step size and order (unchanged part reused) */
    if (h != m->hold) m->ns = 0;
    if (m->ns <= m->kold) m->ns = m->ns + 1;
    const int ns = m->ns, nsp1 = ns + 1;
    if (ns <= k) {
      m->beta[ns - 1] = 1.0;
      m->alpha[ns - 1] = 1.0 / (double)ns;
      temp1 = h * (double)ns;
      m->sig[nsp1 - 1] = 1.0;
      for (int i = nsp1; i <= k; ++i) {
        temp2 = m->psi[i - 2];
        m->psi[i - 2] = temp1;
        m->beta[i - 1] = m->beta[i - 2] * m->psi[i - 2] / temp2;
        temp1 = temp2 + h;
        m->alpha[i - 1] = h / temp1;
        m->sig[i] = (double)i * m->alpha[i - 1] * m->sig[i - 1];
      }
      m->psi[k - 1] = temp1;
      if (ns <= 1) {
        for (int iq = 1; iq <= k; ++iq) {
          m->v[iq - 1] = 1.0 / (double)(iq * (iq + 1));
          m->w[iq - 1] = m->v[iq - 1];
        }
      } else {
        if (m->kold < k) {
          m->v[k - 1] = 1.0 / (double)(k * kp1);
          for (int j = 1; j <= ns - 2; ++j) {
            const int i = k - j;
            m->v[i - 1] = m->v[i - 1] - m->alpha[j] * m->v[i];
          }
        }
        for (int iq = 1; iq <= kp1 - ns; ++iq) {
          m->v[iq - 1] = m->v[iq - 1] - m->alpha[ns - 1] * m->v[iq];
          m->w[iq - 1] = m->v[iq - 1];
        }
        m->g[nsp1 - 1] = m->w[0];
      }
      for (int i = ns + 2; i <= kp1; ++i) {
        for (int iq = 1; iq <= kp2 - i; ++iq) m->w[iq - 1] = m->w[iq - 1] - m->alpha[i - 2] * m->w[iq];
        m->g[i - 1] = m->w[0];
      }
    }
    /* predict: modified differences to the new step, then p = y + h sum g phi */
    for (int i = nsp1; i <= k; ++i)
      for (int l = 0; l < n; ++l) PHI(l, i - 1) = m->beta[i - 1] * PHI(l, i - 1);
    for (int l = 0; l < n; ++l) {
      PHI(l, kp2 - 1) = PHI(l, kp1 - 1);
      PHI(l, kp1 - 1) = 0.0;
      p[l] = 0.0;
    }
    for (int j = 1; j <= k; ++j) {
      const int i = kp1 - j;
      for (int l = 0; l < n; ++l) {
        p[l] = p[l] + m->g[i - 1] * PHI(l, i - 1);
        PHI(l, i - 1) = PHI(l, i - 1) + PHI(l, i);
      }
    }
    if (!m->nornd) {
      for (int l = 0; l < n; ++l) {
        const double tau = h * p[l] - PHI(l, 14);
        p[l] = y[l] + tau;
        PHI(l, 15) = (p[l] - y[l]) - tau;
      }
    } else {
      for (int l = 0; l < n; ++l) p[l] = y[l] + h * p[l];
    }
    const double xold = m->x;
    m->x = m->x + h;
    absh = fabs(h);
    rhs(m, m->x, p, yp);
    /* local errors at orders k-2, k-1, k */
    erkm2 = 0.0; erkm1 = 0.0; erk = 0.0;
    for (int l = 0; l < n; ++l) {
      if (0 < km2) erkm2 = erkm2 + pow((PHI(l, km1 - 1) + yp[l] - PHI(l, 0)) / wt[l], 2);
      if (0 <= km2) erkm1 = erkm1 + pow((PHI(l, k - 1) + yp[l] - PHI(l, 0)) / wt[l], 2);
      erk = erk + pow((yp[l] - PHI(l, 0)) / wt[l], 2);
    }
    if (0 < km2) erkm2 = absh * m->sig[km1 - 1] * kGstr[km2 - 1] * sqrt(erkm2);
    if (0 <= km2) erkm1 = absh * m->sig[k - 1] * kGstr[km1 - 1] * sqrt(erkm1);
    err = absh * sqrt(erk) * (m->g[k - 1] - m->g[kp1 - 1]);
    erk = absh * sqrt(erk) * m->sig[kp1 - 1] * kGstr[k - 1];
    knew = k;
    if (0 < km2) {
      if (dmax(erkm1, erkm2) <= erk) knew = km1;
    } else if (0 == km2) {
      if (erkm1 <= 0.5 * erk) knew = km1;
    }
    if (err <= *eps) break;
    /* rejected: restore the differences and psi, shrink h (order 1 after three failures) */
    m->phase1 = 0;
    m->x = xold;
    for (int i = 1; i <= k; ++i)
      for (int l = 0; l < n; ++l) PHI(l, i - 1) = (PHI(l, i - 1) - PHI(l, i)) / m->beta[i - 1];
    for (int i = 2; i <= k; ++i) m->psi[i - 2] = m->psi[i - 1] - h;
    ifail = ifail + 1;
    temp2 = 0.5;
    if (3 < ifail && p5eps < 0.25 * erk) temp2 = sqrt(p5eps / erk);
    if (3 <= ifail) knew = 1;
    m->h = temp2 * h;
    m->k = knew;
    if (fabs(m->h) < fouru * fabs(m->x)) {
      m->h = fabs(fouru * fabs(m->x)) * sgn(m->h);
      *eps = *eps + *eps;
      return 1;
    }
  }
  /* accepted: correct, evaluate, update the differences */
  const double h = m->h;
  int k = m->k;
  m->kold = k;
  m->hold = h;
  if (!m->nornd) {
    for (int l = 0; l < n; ++l) {
      const double rho = h * m->g[kp1 - 1] * (yp[l] - PHI(l, 0)) - PHI(l, 15);
      y[l] = p[l] + rho;
      PHI(l, 14) = (y[l] - p[l]) - rho;
    }
  } else {
    for (int l = 0; l < n; ++l) y[l] = p[l] + h * m->g[kp1 - 1] * (yp[l] - PHI(l, 0));
  }
  rhs(m, m->x, y, yp);
  for (int l = 0; l < n; ++l) {
    PHI(l, kp1 - 1) = yp[l] - PHI(l, 0);
    PHI(l, kp2 - 1) = PHI(l, kp1 - 1) - PHI(l, kp2 - 1);
  }
  for (int i = 1; i <= k; ++i)
    for (int l = 0; l < n; ++l) PHI(l, i - 1) = PHI(l, i - 1) + PHI(l, kp1 - 1);
  /* order for the next step */
  double erkp1 = 0.0;
  if (knew == km1 || k == 12) m->phase1 = 0;
  if (m->phase1) {
    k = kp1;
    erk = erkp1;
  } else if (knew == km1) {
    k = km1;
    erk = erkm1;
  } else if (kp1 <= m->ns) {
    for (int l = 0; l < n; ++l) erkp1 = erkp1 + pow(PHI(l, kp2 - 1) / wt[l], 2);
    erkp1 = absh * kGstr[kp1 - 1] * sqrt(erkp1);
    if (k == 1) {
      if (erkp1 < 0.5 * erk) { k = kp1; erk = erkp1; }
    } else if (erkm1 <= dmin(erk, erkp1)) {
      k = km1; erk = erkm1;
    } else if (erkp1 < erk && k < 12) {
      k = kp1; erk = erkp1;
    }
  }
  m->k = k;
  /* step size for the next step: double in phase 1, else from the error at the new order */
  double hnew = h + h;
  if (!m->phase1 && p5eps < erk * kTwo[k]) {
    hnew = h;
    if (p5eps < erk) {
      temp2 = (double)(k + 1);
      const double r = pow(p5eps / erk, 1.0 / temp2);
      hnew = absh * dmax(0.5, dmin(0.9, r));
      hnew = fabs(dmax(hnew, fouru * fabs(m->x))) * sgn(h);
    }
  }
  m->h = hnew;
  m->nsteps++;
  return 0;
}

/* DE / ODE: integrate from *t toward tout (beyond it, up to t + 10 (tout - t), then
   interpolate).  iflag as in the reference: 1 on the first call, 2 on normal return and on
   continuation calls, 3 tolerances too small, 4 more than 500 steps, 5 as 4 and stiff,
   6 bad input; negative input iflag stops exactly at tout. */
int adams_ode(adams_mem_t *m, double *y, double *t, double tout, double relerr, double abserr, int *iflag)
{
  const int n = m->n;
  const double fouru = 4.0 * kUround;
  if (n < 1 || *t == tout || relerr < 0.0 || abserr < 0.0) { *iflag = 6; return *iflag; }
  double eps = dmax(relerr, abserr);
  if (eps <= 0.0 || *iflag == 0) { *iflag = 6; return *iflag; }
  const int isn = *iflag < 0 ? -1 : 1;
  *iflag = abs(*iflag);
  if (*iflag != 1) {
    if (*t != m->told || *iflag < 2 || 5 < *iflag) { *iflag = 6; return *iflag; }
  }
  const double del = tout - *t;
  const double absdel = fabs(del);
  const double tend = isn < 0 ? tout : *t + 10.0 * del;
  int nostep = 0, kle4 = 0, stiff = 0;
  const double releps = relerr / eps, abseps = abserr / eps;
  if (*iflag == 1 || m->isnold < 0 || m->delsgn * del <= 0.0) {
    m->start = 1;
    m->x = *t;
    memcpy(m->yy, y, sizeof(double) * n);
    m->delsgn = sgn(del);
    m->h = dmax(fabs(tout - m->x), fouru * fabs(m->x)) * sgn(tout - m->x);
  }
  for (;;) {
    if (absdel <= fabs(m->x - *t)) {
      interpolate(m, tout, y);
      *iflag = 2;
      *t = tout;
      m->told = *t;
      m->isnold = isn;
      return *iflag;
    }
    if (isn <= 0 && fabs(tout - m->x) < fouru * fabs(m->x)) {
      m->h = tout - m->x;
      rhs(m, m->x, m->yy, m->yp);
      for (int l = 0; l < n; ++l) y[l] = m->yy[l] + m->h * m->yp[l];
      *iflag = 2;
      *t = tout;
      m->told = *t;
      m->isnold = isn;
      return *iflag;
    }
    if (ADAMS_MAXNUM <= nostep) {
      *iflag = isn * (stiff ? 5 : 4);
      memcpy(y, m->yy, sizeof(double) * n);
      *t = m->x;
      m->told = *t;
      m->isnold = 1;
      return *iflag;
    }
    m->h = dmin(fabs(m->h), fabs(tend - m->x)) * sgn(m->h);
    for (int l = 0; l < n; ++l) m->wt[l] = releps * fabs(m->yy[l]) + abseps;
    if (adams_step(m, &eps)) {
      *iflag = isn * 3;
      memcpy(y, m->yy, sizeof(double) * n);
      *t = m->x;
      m->told = *t;
      m->isnold = 1;
      return *iflag;
    }
    nostep = nostep + 1;
    kle4 = kle4 + 1;
    if (4 < m->kold) kle4 = 0;
    if (50 <= kle4) stiff = 1;
  }
}

/* ---- drivers of the test problems, the loop of cipa_drug_shago_bench.cpp:119-146 */
static void f_linear(double t, const double *y, double *yp, void *u)
{
  (void)t; (void)u;
  yp[0] = -0.1 * y[0] + 1.0 * y[1];
  yp[1] = -1.0 * y[0] - 0.1 * y[1];
  yp[2] = -0.5 * y[2];
}

static void f_vdp(double t, const double *y, double *yp, void *u)
{
  (void)t; (void)u;
  yp[0] = y[1];
  yp[1] = (1.0 - y[0] * y[0]) * y[1] - y[0];
}

static void f_ord11(double t, const double *y, double *yp, void *u)
{
  double A[ORD11_NA];
  ord11_rates(t, (const double *)u, yp, y, A);
}

/* mode 0 linear, 1 Van der Pol, 2 ORd 2011; in = relerr, abserr, t_end, dt_out, every
   [, conc, bcl, ic50[14]]; out as the golden files (samples of t, y then iflag, n_outputs);
   returns the number of doubles written, or -1 if out_cap is too small */
long oracle_adams_run(int mode, const double *in, double *out, long out_cap, long *nsteps, long *nfe)
{
  double C[ORD11_NC], y[ORD11_NS];
  int n;
  adams_rhs_fn f;
  void *user = NULL;
  if (mode == 0) { n = 3; f = f_linear; y[0] = 1.0; y[1] = 0.0; y[2] = 2.0; }
  else if (mode == 1) { n = 2; f = f_vdp; y[0] = 2.0; y[1] = 0.0; }
  else {
    n = ORD11_NS; f = f_ord11; user = C;
    ord11_init_consts(C, y, 0.0);
    ord11_apply_drug(C, in[5], &in[7]);
    C[140] = in[6];
  }
  const double relerr = in[0], abserr = in[1], t_end = in[2], dt = in[3];
  const int every = (int)in[4];
  adams_mem_t *m = adams_create(n, f, user);
  double t = 0.0, tnext = dt;
  int iflag = 1;
  const long imax = llround(t_end / dt);
  long iout = 0, w = 0;
  while (iout < imax) {
    adams_ode(m, y, &t, tnext, relerr, abserr, &iflag);
    if (iflag != 2) break;
    iout++;
    tnext += dt;
    if (iout % every == 0) {
      if (w + 1 + n > out_cap) { adams_free(m); return -1; }
      out[w++] = t;
      for (int l = 0; l < n; ++l) out[w++] = y[l];
    }
  }
  if (w + 2 > out_cap) { adams_free(m); return -1; }
  out[w++] = iflag;
  out[w++] = (double)iout;
  adams_counts(m, nsteps, nfe, NULL);
  adams_free(m);
  return w;
}
