/*
 * bdf.c -- ORACLE (test infrastructure only): CVODE-style variable-order BDF.
 *
 * The reference integrates every CiPA run with SUNDIALS CVODE configured as
 * CVodeCreate(CV_BDF, CV_NEWTON) + CVDense + CVodeSStolerances + CVodeSetMaxStep(dt),
 * output in CV_NORMAL mode on a dt grid (CEP7Cell/modules/drugs_herg.cpp:151-158,221-229;
 * CEPCell_NEO/modules/cipa_drug_bench.cpp:158-166,213).  CVODE ships in the reference
 * only as a prebuilt archive, so this is a restatement of CVODE's published algorithm
 * (Hindmarsh et al., ACM TOMS 31(3) 2005; CVODE user guide, "Mathematical
 * considerations"): Nordsieck history array, fixed-leading-coefficient BDF of order
 * 1..5, modified Newton with a finite-difference dense Jacobian and LU with partial
 * pivoting, local error test, order/step selection, Jacobian reuse heuristics.
 * CVODE PARITY IS UNPINNED (no reference output pins it); this file is pinned by the
 * textbook/convergence tests in tests/test_oracle_bdf.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define QMAX_ALLOWED 5
#define L_MAX (QMAX_ALLOWED + 1)
/* controller constants (CVODE defaults) */
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 1e-6
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define THRESH 1.5
#define SMALL_NST 10
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define MAXNEF 7
#define MAXNCF 10
#define MAXCOR 3
#define NLSCOEF 0.1
#define CRDOWN 0.3
#define RDIV 2.0
#define MSBP 20
#define DGMAX 0.3
#define MSBJ 50
#define DGMAX_J 0.2
#define HLB_FACTOR 100.0
#define HUB_FACTOR 0.1
#define H_BIAS 0.5
#define MAX_ITERS 4
#define FUZZ_FACTOR 100.0
#define UROUND 2.220446049250313e-16

enum { FIRST_CALL = 0, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { NO_FAILURES = 0, FAIL_BAD_J, FAIL_OTHER };
enum { NLS_OK = 0, NLS_CONV_FAIL = 1, NLS_TRY_AGAIN = 2, NLS_SETUP_FAIL = -1 };

struct bdf_mem {
  int n;
  bdf_rhs_fn f;
  void *user;
  bdf_opts_t o;
  double *zn[L_MAX];
  double *ewt, *y, *acor, *tempv, *ftemp;
  double *J, *savedJ;
  int *piv;
  double tn, h, hprime, eta, hscale, etamax, hu;
  double rl1, gamma, gammap, gamrat, crate, acnrm;
  double tau[L_MAX + 1], tq[6], l[L_MAX];
  int q, qprime, qwait, L, qmax, jcur, indx_acor;
  double saved_tq5;
  long nst, nstlp, nstlj, nfe, nje, nsetups, nni, ncfn, netf;
};

static double wrms(const bdf_mem_t *m, const double *x)
{
  double s = 0.0;
  for (int i = 0; i < m->n; ++i) { const double t = x[i] * m->ewt[i]; s += t * t; }
  return sqrt(s / m->n);
}

static void set_ewt(bdf_mem_t *m, const double *y)
{
  for (int i = 0; i < m->n; ++i) m->ewt[i] = 1.0 / (m->o.rtol * fabs(y[i]) + m->o.atol);
}

bdf_mem_t *bdf_create(int n, bdf_rhs_fn f, void *user, const bdf_opts_t *o)
{
  bdf_mem_t *m = (bdf_mem_t *)calloc(1, sizeof(bdf_mem_t));
  m->n = n; m->f = f; m->user = user; m->o = *o;
  if (m->o.maxord < 1 || m->o.maxord > QMAX_ALLOWED) m->o.maxord = QMAX_ALLOWED;
  if (m->o.max_steps <= 0) m->o.max_steps = 500;
  m->qmax = m->o.maxord;
  for (int j = 0; j < L_MAX; ++j) m->zn[j] = (double *)calloc(n, sizeof(double));
  m->ewt = (double *)calloc(n, sizeof(double));
  m->y = (double *)calloc(n, sizeof(double));
  m->acor = (double *)calloc(n, sizeof(double));
  m->tempv = (double *)calloc(n, sizeof(double));
  m->ftemp = (double *)calloc(n, sizeof(double));
  m->J = (double *)calloc((size_t)n * n, sizeof(double));
  m->savedJ = (double *)calloc((size_t)n * n, sizeof(double));
  m->piv = (int *)calloc(n, sizeof(int));
  return m;
}

void bdf_free(bdf_mem_t *m)
{
  if (!m) return;
  for (int j = 0; j < L_MAX; ++j) free(m->zn[j]);
  free(m->ewt); free(m->y); free(m->acor); free(m->tempv); free(m->ftemp);
  free(m->J); free(m->savedJ); free(m->piv); free(m);
}

int bdf_init(bdf_mem_t *m, double t0, const double *y0)
{
  memcpy(m->zn[0], y0, sizeof(double) * m->n);
  m->f(t0, y0, m->zn[1], m->user);
  m->nfe = 1;
  m->tn = t0;
  m->q = 1; m->L = 2; m->qwait = m->L; m->etamax = ETAMX1;
  m->qprime = 1; m->h = 0.0; m->hprime = 0.0; m->hu = 0.0; m->eta = 1.0;
  m->nst = m->nstlp = m->nstlj = m->nje = m->nsetups = m->nni = m->ncfn = m->netf = 0;
  m->crate = 1.0; m->gamrat = 1.0; m->gammap = 0.0; m->saved_tq5 = 0.0;
  m->indx_acor = m->qmax; m->jcur = 0;
  for (int i = 0; i <= L_MAX; ++i) m->tau[i] = 0.0;
  for (int i = 0; i < 6; ++i) m->tq[i] = 0.0;
  set_ewt(m, y0);
  return 0;
}

/* ---- dense LU with partial pivoting, column-major a[j*n+i] ---- */
static int lu_factor(int n, double *a, int *piv)
{
  for (int k = 0; k < n; ++k) {
    double *ck = a + (size_t)k * n;
    int p = k;
    for (int i = k + 1; i < n; ++i) if (fabs(ck[i]) > fabs(ck[p])) p = i;
    piv[k] = p;
    if (ck[p] == 0.0) return k + 1;
    if (p != k) for (int j = 0; j < n; ++j) { double *cj = a + (size_t)j * n; double t = cj[p]; cj[p] = cj[k]; cj[k] = t; }
    const double mult = 1.0 / ck[k];
    for (int i = k + 1; i < n; ++i) ck[i] *= mult;
    for (int j = k + 1; j < n; ++j) {
      double *cj = a + (size_t)j * n;
      const double akj = cj[k];
      if (akj != 0.0) for (int i = k + 1; i < n; ++i) cj[i] -= akj * ck[i];
    }
  }
  return 0;
}

static void lu_solve(int n, const double *a, const int *piv, double *b)
{
  for (int k = 0; k < n; ++k) { const int p = piv[k]; if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; } }
  for (int k = 0; k < n - 1; ++k) { const double *ck = a + (size_t)k * n; const double bk = b[k]; for (int i = k + 1; i < n; ++i) b[i] -= ck[i] * bk; }
  for (int k = n - 1; k > 0; --k) { const double *ck = a + (size_t)k * n; b[k] /= ck[k]; const double bk = b[k]; for (int i = 0; i < k; ++i) b[i] -= ck[i] * bk; }
  b[0] /= a[0];
}

/* finite-difference Jacobian at (tn, zn[0]) with f(zn[0]) in ftemp (CVODE's DQ Jacobian) */
static void fd_jac(bdf_mem_t *m)
{
  const int n = m->n;
  const double srur = sqrt(UROUND);
  const double fnorm = wrms(m, m->ftemp);
  const double minInc = (fnorm != 0.0) ? (1000.0 * fabs(m->h) * UROUND * n * fnorm) : 1.0;
  double *yy = m->y, *fj = m->tempv;
  memcpy(yy, m->zn[0], sizeof(double) * n);
  for (int j = 0; j < n; ++j) {
    const double ysave = yy[j];
    double inc = srur * fabs(ysave);
    const double alt = minInc / m->ewt[j];
    if (alt > inc) inc = alt;
    yy[j] += inc;
    m->f(m->tn, yy, fj, m->user);
    m->nfe++;
    yy[j] = ysave;
    const double iinv = 1.0 / inc;
    double *col = m->J + (size_t)j * n;
    for (int i = 0; i < n; ++i) col[i] = iinv * fj[i] - iinv * m->ftemp[i];
  }
  m->nje++;
}

static int lsetup(bdf_mem_t *m, int convfail)
{
  const int n = m->n;
  const double dgamma = fabs((m->gamma / m->gammap) - 1.0);
  const int jbad = (m->nst == 0) || (m->nst > m->nstlj + MSBJ) ||
                   ((convfail == FAIL_BAD_J) && (dgamma < DGMAX_J)) || (convfail == FAIL_OTHER);
  if (!jbad) {
    m->jcur = 0;
    memcpy(m->J, m->savedJ, sizeof(double) * n * n);
  } else {
    m->nstlj = m->nst;
    m->jcur = 1;
    fd_jac(m);
    memcpy(m->savedJ, m->J, sizeof(double) * n * n);
  }
  for (size_t k = 0; k < (size_t)n * n; ++k) m->J[k] *= -m->gamma;
  for (int i = 0; i < n; ++i) m->J[(size_t)i * n + i] += 1.0;
  return lu_factor(n, m->J, m->piv) ? 1 : 0;
}

static void lsolve(bdf_mem_t *m, double *b)
{
  lu_solve(m->n, m->J, m->piv, b);
  if (m->gamrat != 1.0) { const double s = 2.0 / (1.0 + m->gamrat); for (int i = 0; i < m->n; ++i) b[i] *= s; }
}

static int newton_iteration(bdf_mem_t *m)
{
  const int n = m->n;
  double del = 0.0, delp = 0.0;
  for (int mm = 0;;) {
    double *b = m->tempv;
    for (int i = 0; i < n; ++i) { const double t = m->rl1 * m->zn[1][i] + m->acor[i]; b[i] = m->gamma * m->ftemp[i] - t; }
    lsolve(m, b);
    m->nni++;
    del = wrms(m, b);
    for (int i = 0; i < n; ++i) { m->acor[i] += b[i]; m->y[i] = m->zn[0][i] + m->acor[i]; }
    if (mm > 0) { const double r = del / delp; m->crate = (CRDOWN * m->crate > r) ? CRDOWN * m->crate : r; }
    const double dcon = del * (m->crate < 1.0 ? m->crate : 1.0) / m->tq[4];
    if (dcon <= 1.0) {
      m->acnrm = (mm == 0) ? del : wrms(m, m->acor);
      m->jcur = 0;
      return NLS_OK;
    }
    ++mm;
    if ((mm == MAXCOR) || ((mm >= 2) && (del > RDIV * delp)))
      return (!m->jcur) ? NLS_TRY_AGAIN : NLS_CONV_FAIL;
    delp = del;
    m->f(m->tn, m->y, m->ftemp, m->user);
    m->nfe++;
  }
}

static int nls(bdf_mem_t *m, int nflag)
{
  int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? NO_FAILURES : FAIL_OTHER;
  int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (m->nst == 0) ||
                  (m->nst >= m->nstlp + MSBP) || (fabs(m->gamrat - 1.0) > DGMAX);
  for (;;) {
    m->f(m->tn, m->zn[0], m->ftemp, m->user);
    m->nfe++;
    if (callSetup) {
      const int ier = lsetup(m, convfail);
      m->nsetups++;
      callSetup = 0;
      m->gamrat = m->crate = 1.0;
      m->gammap = m->gamma;
      m->nstlp = m->nst;
      if (ier) return NLS_CONV_FAIL;
    }
    memset(m->acor, 0, sizeof(double) * m->n);
    memcpy(m->y, m->zn[0], sizeof(double) * m->n);
    const int ier = newton_iteration(m);
    if (ier != NLS_TRY_AGAIN) return ier;
    callSetup = 1;
    convfail = FAIL_BAD_J;
  }
}

static void predict(bdf_mem_t *m)
{
  m->tn += m->h;
  for (int k = 1; k <= m->q; ++k)
    for (int j = m->q; j >= k; --j)
      for (int i = 0; i < m->n; ++i) m->zn[j - 1][i] += m->zn[j][i];
}

static void restore(bdf_mem_t *m, double saved_t)
{
  m->tn = saved_t;
  for (int k = 1; k <= m->q; ++k)
    for (int j = m->q; j >= k; --j)
      for (int i = 0; i < m->n; ++i) m->zn[j - 1][i] -= m->zn[j][i];
}

static void set_tq(bdf_mem_t *m, double hsum, double alpha0, double alpha0_hat, double xi_inv, double xistar_inv)
{
  const int q = m->q;
  const double A1 = 1.0 - alpha0_hat + alpha0;
  const double A2 = 1.0 + q * A1;
  m->tq[2] = fabs(A1 / (alpha0 * A2));
  m->tq[5] = fabs(A2 * xistar_inv / (m->l[q] * xi_inv));
  if (m->qwait == 1) {
    if (q > 1) {
      const double Cc = xistar_inv / m->l[q];
      const double A3 = alpha0 + 1.0 / q;
      const double A4 = alpha0_hat + xi_inv;
      const double Cpinv = (1.0 - A4 + A3) / A3;
      m->tq[1] = fabs(Cc * Cpinv);
    } else {
      m->tq[1] = 1.0;
    }
    hsum += m->tau[q];
    xi_inv = m->h / hsum;
    const double A5 = alpha0 - (1.0 / (q + 1));
    const double A6 = alpha0_hat - xi_inv;
    const double Cppinv = (1.0 - A6 + A5) / A2;
    m->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  m->tq[4] = NLSCOEF / m->tq[2];
}

static void set_bdf(bdf_mem_t *m)
{
  const int q = m->q;
  double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0, hsum = m->h;
  m->l[0] = m->l[1] = 1.0;
  for (int i = 2; i <= q; ++i) m->l[i] = 0.0;
  if (q > 1) {
    for (int j = 2; j < q; ++j) {
      hsum += m->tau[j - 1];
      xi_inv = m->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; --i) m->l[i] += m->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -m->l[1] - alpha0;
    hsum += m->tau[q - 1];
    xi_inv = m->h / hsum;
    alpha0_hat = -m->l[1] - xi_inv;
    for (int i = q; i >= 1; --i) m->l[i] += m->l[i - 1] * xistar_inv;
  }
  set_tq(m, hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
  m->rl1 = 1.0 / m->l[1];
  m->gamma = m->h * m->rl1;
  if (m->nst == 0) m->gammap = m->gamma;
  m->gamrat = (m->nst > 0) ? m->gamma / m->gammap : 1.0;
}

static void rescale(bdf_mem_t *m)
{
  double factor = m->eta;
  for (int j = 1; j <= m->q; ++j) {
    for (int i = 0; i < m->n; ++i) m->zn[j][i] *= factor;
    factor *= m->eta;
  }
  m->h = m->hscale * m->eta;
  m->hscale = m->h;
}

static void increase_order(bdf_mem_t *m)
{
  const int q = m->q, Lc = q + 1;
  double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = m->hscale;
  for (int i = 0; i < L_MAX; ++i) m->l[i] = 0.0;
  m->l[2] = 1.0;
  if (q > 1) {
    for (int j = 1; j < q; ++j) {
      hsum += m->tau[j + 1];
      const double xi = hsum / m->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; --i) m->l[i] = m->l[i] * xiold + m->l[i - 1];
      xiold = xi;
    }
  }
  const double A1 = (-alpha0 - alpha1) / prod;
  for (int i = 0; i < m->n; ++i) m->zn[Lc][i] = A1 * m->zn[m->indx_acor][i];
  for (int j = 2; j <= q; ++j)
    for (int i = 0; i < m->n; ++i) m->zn[j][i] += m->l[j] * m->zn[Lc][i];
}

static void decrease_order(bdf_mem_t *m)
{
  const int q = m->q;
  double hsum = 0.0;
  for (int i = 0; i < L_MAX; ++i) m->l[i] = 0.0;
  m->l[2] = 1.0;
  for (int j = 1; j <= q - 2; ++j) {
    hsum += m->tau[j];
    const double xi = hsum / m->hscale;
    for (int i = j + 2; i >= 2; --i) m->l[i] = m->l[i] * xi + m->l[i - 1];
  }
  for (int j = 2; j < q; ++j)
    for (int i = 0; i < m->n; ++i) m->zn[j][i] -= m->l[j] * m->zn[q][i];
}

static void adjust_order(bdf_mem_t *m, int dq)
{
  if ((m->q == 2) && (dq != 1)) return;
  if (dq == 1) increase_order(m); else if (dq == -1) decrease_order(m);
}

static void set_eta(bdf_mem_t *m)
{
  if (m->eta < THRESH) {
    m->eta = 1.0;
    m->hprime = m->h;
  } else {
    if (m->eta > m->etamax) m->eta = m->etamax;
    const double r = fabs(m->h) * (m->o.hmax > 0 ? 1.0 / m->o.hmax : 0.0) * m->eta;
    m->eta /= (r > 1.0 ? r : 1.0);
    m->hprime = m->h * m->eta;
  }
}

static void prepare_next(bdf_mem_t *m, double dsm)
{
  if (m->etamax == 1.0) {
    if (m->qwait < 2) m->qwait = 2;
    m->qprime = m->q;
    m->hprime = m->h;
    m->eta = 1.0;
    return;
  }
  const double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / m->L) + ADDON);
  if (m->qwait != 0) {
    m->eta = etaq;
    m->qprime = m->q;
    set_eta(m);
    return;
  }
  m->qwait = 2;
  double etaqm1 = 0.0, etaqp1 = 0.0;
  if (m->q > 1) {
    const double ddn = wrms(m, m->zn[m->q]) * m->tq[1];
    etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / m->q) + ADDON);
  }
  if (m->q != m->qmax && m->saved_tq5 != 0.0) {
    const double cquot = (m->tq[5] / m->saved_tq5) * pow(m->h / m->tau[2], (double)m->L);
    for (int i = 0; i < m->n; ++i) m->tempv[i] = m->acor[i] - cquot * m->zn[m->qmax][i];
    const double dup = wrms(m, m->tempv) * m->tq[3];
    etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (m->L + 1)) + ADDON);
  }
  double etam = etaq;
  if (etaqm1 > etam) etam = etaqm1;
  if (etaqp1 > etam) etam = etaqp1;
  if (etam < THRESH) {
    m->eta = 1.0; m->qprime = m->q;
  } else if (etam == etaq) {
    m->eta = etaq; m->qprime = m->q;
  } else if (etam == etaqm1) {
    m->eta = etaqm1; m->qprime = m->q - 1;
  } else {
    m->eta = etaqp1; m->qprime = m->q + 1;
    memcpy(m->zn[m->qmax], m->acor, sizeof(double) * m->n);
  }
  set_eta(m);
}

static void complete_step(bdf_mem_t *m)
{
  m->nst++;
  m->hu = m->h;
  for (int i = m->q; i >= 2; --i) m->tau[i] = m->tau[i - 1];
  if ((m->q == 1) && (m->nst > 1)) m->tau[2] = m->tau[1];
  m->tau[1] = m->h;
  for (int j = 0; j <= m->q; ++j)
    for (int i = 0; i < m->n; ++i) m->zn[j][i] += m->l[j] * m->acor[i];
  m->qwait--;
  if ((m->qwait == 1) && (m->q != m->qmax)) {
    memcpy(m->zn[m->qmax], m->acor, sizeof(double) * m->n);
    m->saved_tq5 = m->tq[5];
    m->indx_acor = m->qmax;
  }
}

/* one internal step; returns 0 or <0 on unrecoverable failure */
static int step(bdf_mem_t *m)
{
  const double saved_t = m->tn;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  double dsm = 0.0;
  if ((m->nst > 0) && (m->hprime != m->h)) {
    if (m->qprime != m->q) {
      adjust_order(m, m->qprime - m->q);
      m->q = m->qprime; m->L = m->q + 1; m->qwait = m->L;
    }
    rescale(m);   /* eta was set with hprime = h * eta by prepare_next */
  }
  for (;;) {
    predict(m);
    set_bdf(m);
    const int ier = nls(m, nflag);
    if (ier != NLS_OK) {
      m->ncfn++;
      restore(m, saved_t);
      ncf++;
      m->etamax = 1.0;
      if ((fabs(m->h) <= m->o.hmin * (1.0 + UROUND)) || (ncf == MAXNCF)) return -4;
      m->eta = (ETACF > m->o.hmin / fabs(m->h)) ? ETACF : m->o.hmin / fabs(m->h);
      nflag = PREV_CONV_FAIL;
      rescale(m);
      continue;
    }
    dsm = m->acnrm * m->tq[2];
    if (dsm <= 1.0) break;
    nef++;
    m->netf++;
    nflag = PREV_ERR_FAIL;
    restore(m, saved_t);
    if ((fabs(m->h) <= m->o.hmin * (1.0 + UROUND)) || (nef == MAXNEF)) return -3;
    m->etamax = 1.0;
    if (nef <= MXNEF1) {
      double eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / m->L) + ADDON);
      const double lo = m->o.hmin / fabs(m->h);
      if (eta < lo) eta = lo;
      if (eta < ETAMIN) eta = ETAMIN;
      if (nef >= SMALL_NEF && eta > ETAMXF) eta = ETAMXF;
      m->eta = eta;
      rescale(m);
      continue;
    }
    if (m->q > 1) {
      m->eta = (ETAMIN > m->o.hmin / fabs(m->h)) ? ETAMIN : m->o.hmin / fabs(m->h);
      adjust_order(m, -1);
      m->L = m->q;
      m->q--;
      m->qwait = m->L;
      rescale(m);
      continue;
    }
    m->eta = (ETAMIN > m->o.hmin / fabs(m->h)) ? ETAMIN : m->o.hmin / fabs(m->h);
    m->h *= m->eta;
    m->hscale = m->h;
    m->qwait = LONG_WAIT;
    m->f(m->tn, m->zn[0], m->tempv, m->user);
    m->nfe++;
    for (int i = 0; i < m->n; ++i) m->zn[1][i] = m->h * m->tempv[i];
  }
  complete_step(m);
  prepare_next(m, dsm);
  m->etamax = (m->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  for (int i = 0; i < m->n; ++i) m->acor[i] *= m->tq[2];
  return 0;
}

static void interpolate(const bdf_mem_t *m, double t, double *yout)
{
  const double s = (t - m->tn) / m->h;
  for (int i = 0; i < m->n; ++i) yout[i] = m->zn[m->q][i];
  for (int j = m->q - 1; j >= 0; --j)
    for (int i = 0; i < m->n; ++i) yout[i] = m->zn[j][i] + s * yout[i];
}

/* initial step size estimate (CVODE's cvHin) */
static double initial_h(bdf_mem_t *m, double tout)
{
  const int n = m->n;
  const double tdist = fabs(tout - m->tn);
  const double tround = UROUND * (fabs(m->tn) > fabs(tout) ? fabs(m->tn) : fabs(tout));
  const double hlb = HLB_FACTOR * tround;
  /* upper bound */
  double hub_inv = 0.0;
  for (int i = 0; i < n; ++i) {
    const double tol = HUB_FACTOR * fabs(m->zn[0][i]) + m->o.atol;
    const double r = fabs(m->zn[1][i]) / tol;
    if (r > hub_inv) hub_inv = r;
  }
  double hub = HUB_FACTOR * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) return hg;
  int hnewOK = 0;
  double hs = hg, hnew = hg;
  for (int count1 = 1; count1 <= MAX_ITERS; ++count1) {
    double *yy = m->y, *ft = m->tempv;
    for (int i = 0; i < n; ++i) yy[i] = hg * m->zn[1][i] + m->zn[0][i];
    m->f(m->tn + hg, yy, ft, m->user);
    m->nfe++;
    for (int i = 0; i < n; ++i) ft[i] = (ft[i] - m->zn[1][i]) * (1.0 / hg);
    const double yddnrm = wrms(m, ft);
    hs = hg;
    if (hnewOK || count1 == MAX_ITERS) { hnew = hg; break; }
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    const double hrat = hnew / hg;
    if ((hrat > 0.5) && (hrat < 2.0)) hnewOK = 1;
    if ((count1 > 1) && (hrat > 2.0)) { hnew = hg; hnewOK = 1; }
    hg = hnew;
  }
  (void)hs;
  double h0 = H_BIAS * hnew;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  return h0;
}

int bdf_advance(bdf_mem_t *m, double tout, double *yout)
{
  if (m->nst == 0 && m->h == 0.0) {
    double h = initial_h(m, tout);
    if (m->o.hmax > 0.0) { const double rh = fabs(h) / m->o.hmax; if (rh > 1.0) h /= rh; }
    if (fabs(h) < m->o.hmin) h *= m->o.hmin / fabs(h);
    m->h = h; m->hscale = h; m->hprime = h;
    for (int i = 0; i < m->n; ++i) m->zn[1][i] *= h;
  } else if ((m->tn - tout) * m->h >= 0.0) {
    interpolate(m, tout, yout);
    return 0;
  }
  for (int nloc = 0;; ++nloc) {
    if (m->nst > 0) set_ewt(m, m->zn[0]);
    if (nloc >= m->o.max_steps) return -1;
    const int k = step(m);
    if (k < 0) return k;
    if ((m->tn - tout) * m->h >= 0.0) { interpolate(m, tout, yout); return 0; }
  }
}

void bdf_get_stats(const bdf_mem_t *m, bdf_stats_t *s)
{
  s->nst = m->nst; s->nfe = m->nfe; s->nsetups = m->nsetups; s->nje = m->nje;
  s->nni = m->nni; s->ncfn = m->ncfn; s->netf = m->netf;
}
double bdf_tn(const bdf_mem_t *m) { return m->tn; }
double bdf_h(const bdf_mem_t *m) { return m->h; }
int bdf_q(const bdf_mem_t *m) { return m->q; }
