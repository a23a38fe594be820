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
// This header holds the per-cell algorithm as operations on one cell (`Lane`): the
// integrator scalars (`Mach`) sit in LDS, the vectors (Nordsieck array, error weights,
// Newton correction, FD Jacobian, factors) sit in HBM as SoA (element k of cell c at
// k*pitch + c).  The kernel in cep_gpu.hip schedules these operations so that every pass
// of a wavefront evaluates one right-hand side per lane.
//
// Newton systems use the Jacobian's structure: 26 Hodgkin-Huxley gates couple only to
// themselves and V, nca to jca/cass, the 9-state Markov IKr block to V, the three charge
// accumulators to nothing downstream, D is constant.  Eliminating gates and the Markov
// block leaves a 12x12 core (CaMKt, 8 concentrations, 2 Jrel, V last).  With V ordered
// last the Markov block (an M-matrix) and the core factor without pivoting; the solution
// is that of the dense system built from the same finite-difference entries (entries
// outside the pattern are exactly zero).
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
// scalar save area between launches
constexpr int SD_N = 40;    // doubles
constexpr int SI_N = 24;    // ints

// controller constants (CVODE defaults)
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1;
constexpr double ETACF = 0.25, ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5;
constexpr int SMALL_NST = 10, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10, MAXNEF = 7, MAXNCF = 10;
constexpr int MAXCOR = 3, MSBP = 20, MSBJ = 50, MAX_ITERS = 4;
constexpr double NLSCOEF = 0.1, CRDOWN = 0.3, RDIV = 2.0, DGMAX = 0.3, DGMAX_J = 0.2;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr double UROUND = 2.220446049250313e-16;

enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };
enum { NLS_OK = 0, NLS_ITER = 1, NLS_TRY_AGAIN = 2, NLS_CONV_FAIL = 3 };
// array work split across the wave: the completion update, the interpolation to an output
// point, the map of the next attempt, the copy of the output to the cell's STATES
enum { SWF_COMPLETE = 1, SWF_INTERP = 2, SWF_ATTEMPT = 4, SWF_YCOPY = 8 };
// upper triangle of the attempt map, row-major: entry (r, c), c >= r
CEP_HD constexpr int map_idx(int r, int c) { return r * 6 - (r * (r - 1)) / 2 + (c - r); }
constexpr int NMAP = 21;
// where a cell's integrator is; PH_F0 .. PH_HIN end a pass with a right-hand-side request
enum {
  PH_FRESH = 0, PH_INIT_F, PH_HIN, PH_START, PH_PREDICT, PH_F0, PH_WAIT_JAC, PH_FACTOR,
  PH_NEWTON, PH_ITER, PH_CONVFAIL, PH_ETEST, PH_RESTART_F, PH_STEP_DONE, PH_DONE
};

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

CEP_HD constexpr bool is_simple_gate(int s) { return (s >= 10 && s <= 32) || (s >= 44 && s <= 46); }
CEP_HD constexpr bool is_core_state(int s) { return (s >= 0 && s <= 9) || s == 47 || s == 48; }

// whether the factorisation or the solve reads J(row, col): the FD Jacobian's structural pattern
// as the structured LU uses it (gate rows: diagonal and V; nca: nca, jca, cass; the Markov block
// with its V column; core and accumulator rows: the core columns and their coupled gates, nca
// and O).  The finite-difference columns store only these entries.
CEP_HD constexpr bool jac_read(int row, int col)
{
  if (is_simple_gate(row)) return col == row || col == 0;
  if (row == 33) return col == 33 || col == 30 || col == 2;
  if (row >= 34 && row <= 42) return (col >= 34 && col <= 42) || col == 0;
  if (is_core_state(row) || (row >= 49 && row <= 51))
    return is_core_state(col) || (is_simple_gate(col) && couples(row, col)) || (col == 33 && couples(row, 33)) ||
           (col == 38 && couples(row, 38));
  return false;
}

// entries fetch_couplings gathers for one row: 12 core columns (optional), the coupled simple
// gates, nca and O
CEP_HD constexpr int ncouplings(int row, bool with_core)
{
  int n = with_core ? 12 : 0;
  for (int k = 0; k < NG; ++k) n += couples(row, kSimpleGates[k]) ? 1 : 0;
  return n + (couples(row, 33) ? 1 : 0) + (couples(row, 38) ? 1 : 0);
}
constexpr int NCPL_MAX = 40;
static_assert(ncouplings(0, true) == NCPL_MAX && ncouplings(49, true) <= NCPL_MAX,
              "coupling batch size");

struct Buffers {
  int ncell, pitch;
  double *y;        // [52] state at the last output point (CellML STATES)
  // integrator vectors, in blocks of CPB cells
  double *zn;       // [6*52] Nordsieck array
  double *ewt, *acor, *ft;   // [52]: error weights, Newton correction, f at the current iterate
  double *jd;       // [48*52] FD Jacobian, column slot-major
  double *fac;      // [NFAC]
  double *sd;       // [SD_N] scalar save
  int *si;          // [SI_N]
  double *prm;      // [ord::NP] per-cell parameters
  double *bmd;      // [BM_TOTAL_D]
  int *bmi;         // [BM_I_N]
  double *ct, *cv;  // Ca trace [cap]
  double *vtrace;   // optional Vm trace of the last beat [tcap], or null
  TsView ts;        // optional time series of the selected beat (ts.t == null: off)
  int tcap;
  long long *stats; // [8] nst nfe nje nsetups nni ncfn netf nout
  int *status;
  double dt, rtol, atol, uround_fd;
  int maxord, max_steps;
  int cap;
  const int *gorder;  // [groups] cell group of each workgroup (ordered beat launches), or null
  const int *win;     // per-cell first biomarker window beat, -1: the batch's window
};

// integrator scalars of one cell (LDS-resident during a launch)
struct Mach {
  double tn, h, hprime, eta, hscale, etamax, rl1, gamma, gammap, gamrat, crate, acnrm, del, delp;
  double saved_t, tq1, tq2, tq3, tq4, tq5, saved_tq5, tout, hg, hub, hlb, hnew, mininc;
  double tau[7], l[6];
  int q, qprime, qwait, jcur, phase, callSetup, convfail, nflag, ncf, nef, mnewt;
  int iout, iout_end, status, hin_count, hin_ok, nstloc, jend;
  int sw_ops, sw_dq, sw_q1;   // Nordsieck-array work pending for the next attempt
  // array work of this pass done by the whole wave (SWF_* flags) and its parameters
  int wsw, c_save, c_dn, c_up, i_q5;
  double c_cquot, c_dsm, i_s;
  long long nst, nstlp, nstlj, nfe, nje, nsetups, nni, ncfn, netf;
};

// The integrator's per-cell vectors are stored in blocks of CPB cells, one block per
// workgroup: element k of the cell in block slot `ln` is at base[k*CPB + ln].  They are
// accessed with buffer instructions through one resource descriptor per vector (scalar
// registers), the slot offset ln*8 (one vector register) and a compile-time byte offset
// split into the instruction's immediate (< 4 KiB) and a scalar constant: no per-lane
// address arithmetic, and accesses outside a block's range read 0 / are dropped.
constexpr int CPB = 16;

typedef __amdgpu_buffer_rsrc_t Rsrc;
CEP_D Rsrc make_rsrc(const double *base, unsigned bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000);
}

// one double of a blocked vector, readable and assignable like a reference
struct BElem {
  Rsrc r;
  unsigned voff;
  int byte;
  CEP_D operator double() const
  {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, byte, 0));
  }
  CEP_D const BElem &operator=(double v) const
  {
    typedef __attribute__((ext_vector_type(2))) unsigned u2;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, byte, 0);
    return *this;
  }
};

// a lane's 52-double staging vector in LDS, laid out [component][lane]
struct LVec {
  double *p;
  CEP_D double &operator[](int i) const { return p[i * 64]; }
};

struct BlockBufs {
  Rsrc zn, ewt, acor, ft, jd, fac;
};

// one cell's view: buffers (kernel arguments, wave-uniform), its LDS scalars, the block
// of integrator vectors, its slot in the block, its index and the pitch of the per-cell
// (SoA: element k of cell c at k*p + c) buffers shared with the host.  The kernel passes
// the pitch re-read every pass, which keeps the compiler from hoisting loop-invariant
// addresses of those buffers out of the integration loop.
template <typename R>
struct Lane {
  const Buffers &B;
  Mach &m;
  const BlockBufs &V;
  unsigned ln;
  unsigned c;
  int p;

  // powers of the step-size and order heuristics (the eta factors, the order-change
  // quotient), in the precision of the rates: they only steer the choice of step and order,
  // and an FP64 pow is a long serial sequence on the integrator lanes
  CEP_D static double hpow(double x, double e)
  {
    if constexpr (sizeof(R) == 4) return (double)__powf((float)x, (float)e);
    else return pow(x, e);
  }

  // (the mask tells the compiler the slot offset is small, so constants fold into the
  // instruction's immediate offset)
  CEP_D BElem blk(Rsrc r, int k) const { return BElem{r, (ln & 63u) * 8u, k * CPB * 8}; }
  CEP_D BElem ZN(int j, int i) const { return blk(V.zn, j * N + i); }
  CEP_D BElem EWT(int i) const { return blk(V.ewt, i); }
  CEP_D BElem ACOR(int i) const { return blk(V.acor, i); }
  CEP_D BElem FT(int i) const { return blk(V.ft, i); }
  CEP_D BElem JD(int slot, int row) const { return blk(V.jd, slot * N + row); }
  CEP_D double J(int row, int col) const { return JD(jslot(col), row); }
  CEP_D BElem FAC(int k) const { return blk(V.fac, k); }
  // the same elements for an index that differs between lanes (a Jacobian column slot, a
  // perturbed component): the index goes into the per-lane offset.  As a scalar offset a
  // divergent index would make the compiler loop over its distinct values.
  CEP_D BElem blk_lane(Rsrc r, unsigned k, int kc) const
  {
    return BElem{r, (ln & 63u) * 8u + k * (CPB * 8u), kc * CPB * 8};
  }
  CEP_D BElem ZN_lane(int j, unsigned i) const { return blk_lane(V.zn, i, j * N); }
  CEP_D BElem EWT_lane(unsigned i) const { return blk_lane(V.ewt, i, 0); }
  CEP_D BElem JD_lane(unsigned slot, int row) const { return blk_lane(V.jd, slot * N, row); }
  CEP_D double &Y(int i) const { return B.y[i * p + c]; }

  // ---------------- persistence between launches
  CEP_D void load()
  {
    const double *d = B.sd + c;
    double *dv[SD_N] = { &m.tn, &m.h, &m.hprime, &m.eta, &m.hscale, &m.etamax, &m.rl1, &m.gamma,
                         &m.gammap, &m.gamrat, &m.crate, &m.acnrm, &m.del, &m.delp, &m.saved_t,
                         &m.tq1, &m.tq2, &m.tq3, &m.tq4, &m.tq5, &m.saved_tq5, &m.tout, &m.hg,
                         &m.hub, &m.hlb, &m.hnew, &m.mininc, &m.tau[0], &m.tau[1], &m.tau[2],
                         &m.tau[3], &m.tau[4], &m.tau[5], &m.tau[6], &m.l[0], &m.l[1], &m.l[2],
                         &m.l[3], &m.l[4], &m.l[5] };
#pragma unroll
    for (int k = 0; k < SD_N; ++k) *dv[k] = d[k * p];
    const int *s = B.si + c;
    int *iv[14] = { &m.q, &m.qprime, &m.qwait, &m.jcur, &m.phase, &m.callSetup, &m.convfail,
                    &m.nflag, &m.ncf, &m.nef, &m.mnewt, &m.iout, &m.hin_count, &m.hin_ok };
#pragma unroll
    for (int k = 0; k < 14; ++k) *iv[k] = s[k * p];
    m.nst = s[14 * p]; m.nstlp = s[15 * p]; m.nstlj = s[16 * p];
    m.status = B.status[c];
    const long long *st = B.stats + c;
    m.nfe = st[1 * p]; m.nje = st[2 * p]; m.nsetups = st[3 * p]; m.nni = st[4 * p];
    m.ncfn = st[5 * p]; m.netf = st[6 * p];
  }
  CEP_D void store() const
  {
    double *d = B.sd + c;
    const double dv[SD_N] = { m.tn, m.h, m.hprime, m.eta, m.hscale, m.etamax, m.rl1, m.gamma,
                              m.gammap, m.gamrat, m.crate, m.acnrm, m.del, m.delp, m.saved_t,
                              m.tq1, m.tq2, m.tq3, m.tq4, m.tq5, m.saved_tq5, m.tout, m.hg, m.hub,
                              m.hlb, m.hnew, m.mininc, m.tau[0], m.tau[1], m.tau[2], m.tau[3],
                              m.tau[4], m.tau[5], m.tau[6], m.l[0], m.l[1], m.l[2], m.l[3], m.l[4],
                              m.l[5] };
#pragma unroll
    for (int k = 0; k < SD_N; ++k) d[k * p] = dv[k];
    int *s = B.si + c;
    const int iv[14] = { m.q, m.qprime, m.qwait, m.jcur, m.phase, m.callSetup, m.convfail,
                         m.nflag, m.ncf, m.nef, m.mnewt, m.iout, m.hin_count, m.hin_ok };
#pragma unroll
    for (int k = 0; k < 14; ++k) s[k * p] = iv[k];
    s[14 * p] = (int)m.nst; s[15 * p] = (int)m.nstlp; s[16 * p] = (int)m.nstlj;
    B.status[c] = m.status;
    long long *st = B.stats + c;
    st[0] = m.nst; st[1 * p] = m.nfe; st[2 * p] = m.nje; st[3 * p] = m.nsetups; st[4 * p] = m.nni;
    st[5 * p] = m.ncfn; st[6 * p] = m.netf; st[7 * p] = m.iout;
  }

  // ---------------- vector sweeps over the 52 components.  A sweep issues the loads of a
  // chunk of CH components together, computes in registers, then stores the chunk: the
  // compiler cannot move a load above a store to the same buffers, so per-component
  // load/store interleaving would serialise every load on memory latency.
  static constexpr int CH = 13;
  static_assert(N % CH == 0, "whole chunks");

  // CVODE's weights from zn[0]: 1 / (rtol |y| + atol)
  CEP_D void set_ewt()
  {
    const double rt = B.rtol, at_ = B.atol;
    double w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = 1.0 / (rt * fabs(ZN(0, i)) + at_);
#pragma unroll
    for (int i = 0; i < N; ++i) EWT(i) = w[i];
  }

  // Pascal-triangle update of one component's Nordsieck column (CVODE's prediction,
  // sign +1, and its undo, sign -1), order q
  CEP_D static void pascal_reg(double *z, int q, int sign)
  {
#pragma unroll
    for (int k = 1; k <= QMAX; ++k)
#pragma unroll
      for (int j = QMAX; j >= k; --j)
        if (k <= q && j <= q) z[j - 1] = sign > 0 ? z[j - 1] + z[j] : z[j - 1] - z[j];
  }

  // SW_RESTART: after the restore, evaluate f(tn, zn[0]) for zn[1] instead of predicting
  enum { SW_RESTORE = 1, SW_INC = 2, SW_DEC = 4, SW_RESCALE = 8, SW_PREDICT = 16, SW_RESTART = 32 };
  // Everything done to the Nordsieck array between two nonlinear solves, in CVODE's order:
  // undo a failed prediction (order q0), change the order (cvIncreaseBDF / cvDecreaseBDF
  // with coefficients ll and a1, order q0), rescale row j by eta^j (order q1), predict
  // (order q1).  All of it is linear in a component's column (zn[0..5][i]), so the
  // operations are first composed into one upper-triangular 6x6 map M (build_map: applied
  // to the unit columns, in registers), then the second wave sweep (output_attempt_part)
  // computes M z for the 52 components with no order-dependent branches.  Rows above the new
  // order, except the saved correction in row QMAX, leave the map as zeros: the array keeps
  // rows q+1..QMAX-1 at zero, which the interpolation and the other sweeps rely on.  The
  // staging vector receives the new zn[0].
  CEP_D static void build_map(int ops, int q0, int q1, const double *ll, double a1, double eta,
                              double (&M)[QMAX + 1][QMAX + 1])
  {
    double fac[QMAX + 1];
    fac[0] = 1.0;
#pragma unroll
    for (int j = 1; j <= QMAX; ++j) fac[j] = fac[j - 1] * eta;
#pragma unroll
    for (int col = 0; col <= QMAX; ++col) {
      double zz[QMAX + 1];
#pragma unroll
      for (int r = 0; r <= QMAX; ++r) zz[r] = (r == col) ? 1.0 : 0.0;
      if (ops & SW_RESTORE) pascal_reg(zz, q0, -1);
      if (ops & SW_INC) {
        const double zl = a1 * zz[QMAX];
#pragma unroll
        for (int j = 2; j <= QMAX; ++j) if (j <= q0) zz[j] += ll[j] * zl;
#pragma unroll
        for (int j = 2; j <= QMAX; ++j) if (j == q0 + 1) zz[j] = zl;
      }
      if (ops & SW_DEC) {
        double zq = 0.0;
#pragma unroll
        for (int j = 2; j <= QMAX; ++j) if (j == q0) zq = zz[j];
#pragma unroll
        for (int j = 2; j < QMAX; ++j) if (j < q0) zz[j] -= ll[j] * zq;
      }
      if (ops & SW_RESCALE) {
#pragma unroll
        for (int j = 1; j <= QMAX; ++j) if (j <= q1) zz[j] *= fac[j];
      }
      if (ops & SW_PREDICT) pascal_reg(zz, q1, 1);
#pragma unroll
      for (int r = 0; r < QMAX; ++r) if (r > q1) zz[r] = 0.0;
#pragma unroll
      for (int r = 0; r <= QMAX; ++r) M[r][col] = zz[r];
    }
  }

  CEP_D void set_bdf()
  {
    const int q = m.q;
    double l[QMAX + 1];
    double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0, hsum = m.h;
    l[0] = l[1] = 1.0;
#pragma unroll
    for (int i = 2; i <= QMAX; ++i) l[i] = 0.0;
    if (q > 1) {
      for (int j = 2; j < q; ++j) {
        hsum += m.tau[j - 1];
        xi_inv = m.h / hsum;
        alpha0 -= 1.0 / j;
#pragma unroll
        for (int i = QMAX; i >= 1; --i) if (i <= j) l[i] += l[i - 1] * xi_inv;
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += m.tau[q - 1];
      xi_inv = m.h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = QMAX; i >= 1; --i) if (i <= q) l[i] += l[i - 1] * xistar_inv;
    }
    double lq = l[1];
#pragma unroll
    for (int i = 2; i <= QMAX; ++i) if (i == q) lq = l[i];
#pragma unroll
    for (int i = 0; i <= QMAX; ++i) m.l[i] = l[i];
    // error-test / order-change coefficients
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    m.tq2 = fabs(A1 / (alpha0 * A2));
    m.tq5 = fabs(A2 * xistar_inv / (lq * xi_inv));
    if (m.qwait == 1) {
      if (q > 1) {
        const double Cc = xistar_inv / lq;
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        m.tq1 = fabs(Cc * Cpinv);
      } else {
        m.tq1 = 1.0;
      }
      const double hsum2 = hsum + m.tau[q];
      const double xi2 = m.h / hsum2;
      const double A5 = alpha0 - (1.0 / (q + 1));
      const double A6 = alpha0_hat - xi2;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      m.tq3 = fabs(Cppinv / (xi2 * (q + 2) * A5));
    }
    m.tq4 = NLSCOEF / m.tq2;
    m.rl1 = 1.0 / l[1];
    m.gamma = m.h * m.rl1;
    if (m.nst == 0) m.gammap = m.gamma;
    m.gamrat = (m.nst > 0) ? m.gamma / m.gammap : 1.0;
  }

  // cvIncreaseBDF coefficients at order q: zn[q+1] = a1 zn[qmax], zn[j] += ll[j] zn[q+1]
  CEP_D double inc_coeffs(double *ll) const
  {
    const int q = m.q;
#pragma unroll
    for (int i = 0; i <= QMAX; ++i) ll[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = m.hscale;
    ll[2] = 1.0;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += m.tau[j + 1];
        const double xi = hsum / m.hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
#pragma unroll
        for (int i = QMAX; i >= 2; --i) if (i <= j + 2) ll[i] = ll[i] * xiold + ll[i - 1];
        xiold = xi;
      }
    }
    return (-alpha0 - alpha1) / prod;
  }
  // cvDecreaseBDF coefficients at order q: zn[j] -= ll[j] zn[q], 2 <= j < q
  CEP_D void dec_coeffs(double *ll) const
  {
    const int q = m.q;
#pragma unroll
    for (int i = 0; i <= QMAX; ++i) ll[i] = 0.0;
    ll[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += m.tau[j];
      const double xi = hsum / m.hscale;
#pragma unroll
      for (int i = QMAX; i >= 2; --i) if (i <= j + 2) ll[i] = ll[i] * xi + ll[i - 1];
    }
  }
  // order change dq at the current order (CVODE's cvAdjustOrder: no decrease below 2),
  // folded into `ops`/`ll`/`a1` for build_map
  CEP_D int order_ops(int dq, double *ll, double &a1) const
  {
    a1 = 0.0;
    if ((m.q == 2) && (dq != 1)) return 0;
    if (dq == 1) { a1 = inc_coeffs(ll); return SW_INC; }
    if (dq == -1) { dec_coeffs(ll); return SW_DEC; }
    return 0;
  }

  // ---------------- linear algebra: structured factorisation of M = I - gamma*J.
  // Each section fetches its Jacobian entries in batches (fetch / fetch_couplings below),
  // computes in registers and stores its factors.
  CEP_D bool factor()
  {
    const double g = m.gamma;
    // Markov 9x9 LU (no pivoting: column diagonally dominant M-matrix), w = M^-1 (-g J_mV)
    double A[9][9], w[9];
    {
      double jm[45], jw[9];
#pragma unroll
      for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) jm[i * 9 + j] = J(34 + i, 34 + j);
#pragma unroll
      for (int i = 0; i < 9; ++i) jw[i] = J(34 + i, 0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 5; ++i) {
#pragma unroll
        for (int j = 0; j < 9; ++j) A[i][j] = (i == j ? 1.0 : 0.0) - g * jm[i * 9 + j];
      }
#pragma unroll
      for (int i = 0; i < 9; ++i) w[i] = -g * jw[i];
#pragma unroll
      for (int i = 5; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) jm[(i - 5) * 9 + j] = J(34 + i, 34 + j);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 5; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) A[i][j] = (i == j ? 1.0 : 0.0) - g * jm[(i - 5) * 9 + j];
    }
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
    const double w4 = w[4];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
#pragma unroll
      for (int j = 0; j < 9; ++j) FAC(F_MLU + i * 9 + j) = A[i][j];
      FAC(F_W + i) = w[i];
    }
    // simple gates: x_g = r_g b_g - c_g x_V; nca: row {cass, jca, nca}
    double cg[NG], cnV, cnC;
    {
      double jd[NG], jv[NG], jn[3];
#pragma unroll
      for (int k = 0; k < NG; ++k) {
        jd[k] = J(kSimpleGates[k], kSimpleGates[k]);
        jv[k] = J(kSimpleGates[k], 0);
      }
      jn[0] = J(33, 33);
      jn[1] = J(33, 30);
      jn[2] = J(33, 2);
      __builtin_amdgcn_sched_barrier(0);
      double rg[NG];
#pragma unroll
      for (int k = 0; k < NG; ++k) {
        const double r = 1.0 / (1.0 - g * jd[k]);
        rg[k] = r;
        cg[k] = (-g * jv[k]) * r;
      }
      const double rn = 1.0 / (1.0 - g * jn[0]);
      const double e_nj = rn * (-g * jn[1]);
      cnV = -e_nj * cg[20];                      // jca is simple gate #20
      cnC = rn * (-g * jn[2]);
#pragma unroll
      for (int k = 0; k < NG; ++k) { FAC(F_RG + k) = rg[k]; FAC(F_CG + k) = cg[k]; }
      FAC(F_NCA + 0) = rn;
      FAC(F_NCA + 1) = e_nj;
      FAC(F_NCA + 2) = cnV;
      FAC(F_NCA + 3) = cnC;
    }
    // core Schur complement (12x12, V last), one row's entries per batch
    double S[12][12];
#pragma unroll
    for (int a = 0; a < 12; ++a) {
      const int row = kCore[a];
      double jr[NCPL_MAX];
      fetch_couplings(jr, row, true);
      int n = 0;
#pragma unroll
      for (int b = 0; b < 12; ++b) S[a][b] = (a == b ? 1.0 : 0.0) - g * jr[n++];
      double sv = 0.0;
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(row, kSimpleGates[k])) sv += (-g * jr[n++]) * cg[k];
      if (couples(row, 33)) {
        const double mn = -g * jr[n++];
        sv += mn * cnV;
        S[a][CORE_CASS] -= mn * cnC;
      }
      if (couples(row, 38)) sv += (-g * jr[n++]) * w4;
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

  // K consecutive elements of a blocked vector, loaded together: the scheduling barrier
  // keeps the compiler from sinking each load to its first use, which would put one memory
  // round trip per element on the critical path.
  template <int K>
  CEP_D void fetch(double (&r)[K], Rsrc rs, int k0) const
  {
#pragma unroll
    for (int k = 0; k < K; ++k) r[k] = blk(rs, k0 + k);
    __builtin_amdgcn_sched_barrier(0);
  }
  // the Jacobian entries J(row, col) of the solve's coupling terms for one row, in the order
  // the solve uses them (simple gates, nca, O); returns their number
  CEP_D int fetch_couplings(double (&r)[NCPL_MAX], int row, bool with_core) const
  {
    int n = 0;
    if (with_core) {
#pragma unroll
      for (int a = 0; a < 12; ++a) r[n++] = J(row, kCore[a]);
    }
#pragma unroll
    for (int k = 0; k < NG; ++k)
      if (couples(row, kSimpleGates[k])) r[n++] = J(row, kSimpleGates[k]);
    if (couples(row, 33)) r[n++] = J(row, 33);
    if (couples(row, 38)) r[n++] = J(row, 38);
    __builtin_amdgcn_sched_barrier(0);
    return n;
  }

  // number of coupling entries (simple gates, nca, O) of the core rows kCore[A0..A1)
  template <int A0, int A1>
  CEP_HD static constexpr int ncoupl_rows()
  {
    int n = 0;
    for (int a = A0; a < A1; ++a) n += ncouplings(kCore[a], false);
    return n;
  }
  // bk[a] += g * J(row, gate) * beta_gate + ... for the core rows kCore[A0..A1): every entry of the
  // rows fetched in one batch, then the sums in the order of the rows
  template <int A0, int A1>
  CEP_D void core_couplings(double (&bk)[12], const double (&beta)[NG], double beta_n, double bO, double g) const
  {
    constexpr int NB = ncoupl_rows<A0, A1>();
    static_assert(NB > 0 && NB <= NCPL_MAX, "coupling batch size");
    double jr[NB];
    int n = 0;
#pragma unroll
    for (int a = A0; a < A1; ++a) {
      const int row = kCore[a];
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(row, kSimpleGates[k])) jr[n++] = J(row, kSimpleGates[k]);
      if (couples(row, 33)) jr[n++] = J(row, 33);
      if (couples(row, 38)) jr[n++] = J(row, 38);
    }
    __builtin_amdgcn_sched_barrier(0);
    n = 0;
#pragma unroll
    for (int a = A0; a < A1; ++a) {
      const int row = kCore[a];
      double s = bk[a];
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(row, kSimpleGates[k])) s += g * jr[n++] * beta[k];
      if (couples(row, 33)) s += g * jr[n++] * beta_n;
      if (couples(row, 38)) s += g * jr[n++] * bO;
      bk[a] = s;
    }
  }
  // accumulator rows R0..R1-1 (qnet, INaL and ICaL AUC): x_r = b_r + g * J(r, :) x, the rows'
  // entries fetched in one batch
  template <int R0, int R1, typename BV>
  CEP_D void acc_rows(const BV &b, double g) const
  {
    constexpr int NB = [] { int n = 0; for (int r = R0; r < R1; ++r) n += ncouplings(r, true); return n; }();
    static_assert(NB <= NCPL_MAX, "accumulator batch size");
    double jr[NB];
    int n = 0;
#pragma unroll
    for (int r = R0; r < R1; ++r) {
#pragma unroll
      for (int a = 0; a < 12; ++a) jr[n++] = J(r, kCore[a]);
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(r, kSimpleGates[k])) jr[n++] = J(r, kSimpleGates[k]);
      if (couples(r, 33)) jr[n++] = J(r, 33);
      if (couples(r, 38)) jr[n++] = J(r, 38);
    }
    __builtin_amdgcn_sched_barrier(0);
    n = 0;
#pragma unroll
    for (int r = R0; r < R1; ++r) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 12; ++a) s += jr[n++] * b[kCore[a]];
#pragma unroll
      for (int k = 0; k < NG; ++k)
        if (couples(r, kSimpleGates[k])) s += jr[n++] * b[kSimpleGates[k]];
      if (couples(r, 33)) s += jr[n++] * b[33];
      if (couples(r, 38)) s += jr[n++] * b[38];
      b[r] += g * s;
    }
  }

  // solve M x = b in place (b: a lane vector, here the LDS staging vector), then CVODE's
  // 2/(1+gamrat) scaling.  Factors and Jacobian entries are fetched in batches.
  template <typename BV>
  CEP_D void solve(const BV &b) const
  {
    const double g = m.gammap;
    double beta[NG], beta_n;
    {
      double rg[NG], nc[2];
      fetch(rg, V.fac, F_RG);
      fetch(nc, V.fac, F_NCA);
#pragma unroll
      for (int k = 0; k < NG; ++k) beta[k] = rg[k] * b[kSimpleGates[k]];
      beta_n = nc[0] * b[33] - nc[1] * beta[20];
    }
    double bm[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) bm[i] = b[34 + i];
    {
      // Markov LU: rows 1-8 of L, then rows 8..0 of U (with the diagonal)
      double lo[36];
      int n = 0;
#pragma unroll
      for (int i = 1; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) lo[n++] = FAC(F_MLU + i * 9 + j);
      __builtin_amdgcn_sched_barrier(0);
      n = 0;
#pragma unroll
      for (int i = 1; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) bm[i] -= lo[n++] * bm[j];
    }
    {
      double up[45];
      int n = 0;
#pragma unroll
      for (int i = 8; i >= 0; --i)
#pragma unroll
        for (int j = i; j < 9; ++j) up[n++] = FAC(F_MLU + i * 9 + j);
      __builtin_amdgcn_sched_barrier(0);
      n = 0;
#pragma unroll
      for (int i = 8; i >= 0; --i) {
        const double d = up[n++];
#pragma unroll
        for (int j = i + 1; j < 9; ++j) bm[i] -= up[n++] * bm[j];
        bm[i] /= d;
      }
    }
    // the core rows' right-hand sides: the coupling entries of two rows per batch (one memory
    // round trip each; CaMKt and the SR / Ca_i rows have none)
    double bk[12];
#pragma unroll
    for (int a = 0; a < 12; ++a) bk[a] = b[kCore[a]];
    core_couplings<0, 3>(bk, beta, beta_n, bm[4], g);    // CaMKt, cass, nai
    core_couplings<3, 5>(bk, beta, beta_n, bm[4], g);    // nass, ki
    core_couplings<5, 10>(bk, beta, beta_n, bm[4], g);   // kss, cansr, cajsr, cai, Jrel np
    core_couplings<10, 12>(bk, beta, beta_n, bm[4], g);  // Jrel p, V
    {
      // core LU: rows 1-11 of L in two batches, then U rows 11..0 in two
      double lo[38];
      int n = 0;
#pragma unroll
      for (int i = 1; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) lo[n++] = FAC(F_SLU + i * 12 + j);
      __builtin_amdgcn_sched_barrier(0);
      n = 0;
#pragma unroll
      for (int i = 1; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) bk[i] -= lo[n++] * bk[j];
      n = 0;
#pragma unroll
      for (int i = 8; i < 12; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) lo[n++] = FAC(F_SLU + i * 12 + j);
      __builtin_amdgcn_sched_barrier(0);
      n = 0;
#pragma unroll
      for (int i = 8; i < 12; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) bk[i] -= lo[n++] * bk[j];
    }
    {
      double up[42];
      int n = 0;
#pragma unroll
      for (int i = 11; i >= 4; --i)
#pragma unroll
        for (int j = i; j < 12; ++j) up[n++] = FAC(F_SLU + i * 12 + j);
      __builtin_amdgcn_sched_barrier(0);
      n = 0;
#pragma unroll
      for (int i = 11; i >= 4; --i) {
        const double d = up[n++];
#pragma unroll
        for (int j = i + 1; j < 12; ++j) bk[i] -= up[n++] * bk[j];
        bk[i] /= d;
      }
      n = 0;
#pragma unroll
      for (int i = 3; i >= 0; --i)
#pragma unroll
        for (int j = i; j < 12; ++j) up[n++] = FAC(F_SLU + i * 12 + j);
      __builtin_amdgcn_sched_barrier(0);
      n = 0;
#pragma unroll
      for (int i = 3; i >= 0; --i) {
        const double d = up[n++];
#pragma unroll
        for (int j = i + 1; j < 12; ++j) bk[i] -= up[n++] * bk[j];
        bk[i] /= d;
      }
    }
    const double xV = bk[CORE_V], xC = bk[CORE_CASS];
#pragma unroll
    for (int a = 0; a < 12; ++a) b[kCore[a]] = bk[a];
    {
      double cg[NG], nc[2], w[9];
      fetch(cg, V.fac, F_CG);
      fetch(nc, V.fac, F_NCA + 2);
      fetch(w, V.fac, F_W);
#pragma unroll
      for (int k = 0; k < NG; ++k) b[kSimpleGates[k]] = beta[k] - cg[k] * xV;
      b[33] = beta_n - nc[0] * xV - nc[1] * xC;
#pragma unroll
      for (int i = 0; i < 9; ++i) b[34 + i] = bm[i] - w[i] * xV;
    }
    // accumulators: x = b + gamma * J(acc, :) x, row 49 then rows 50 and 51 in one batch
    acc_rows<49, 50>(b, g);
    acc_rows<50, 52>(b, g);
    if (m.gamrat != 1.0) {
      const double sc = 2.0 / (1.0 + m.gamrat);
#pragma unroll
      for (int i = 0; i < N; ++i) b[i] *= sc;
    }
  }

  // One modified-Newton iteration.  `y` (the lane's staging vector) holds f(tn, y_iterate)
  // on entry; the Newton right-hand side and the solve run in place there, and on exit it
  // holds the next iterate.  Returns NLS_OK, NLS_ITER (evaluate f(y) and call again),
  // NLS_TRY_AGAIN or NLS_CONV_FAIL (CVODE's cvNewtonIteration).
  template <typename YV>
  CEP_D int newton(const YV &y)
  {
    const int mm = m.mnewt;
    const double gm = m.gamma, r1 = m.rl1;
    // half the components per chunk: two memory round trips per loop (chunks of 13 took four)
    constexpr int NCH = N / 2;
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += NCH) {
      double z1[NCH], a[NCH];
      fetch(z1, V.zn, N + c0);
      if (mm > 0) fetch(a, V.acor, c0);
      else {
#pragma unroll
        for (int k = 0; k < NCH; ++k) a[k] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < NCH; ++k) y[c0 + k] = gm * y[c0 + k] - (r1 * z1[k] + a[k]);
    }
    solve(y);
    m.nni++;
    double s2 = 0.0, sa = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += NCH) {
      double a[NCH], z[NCH], w[NCH];
#pragma unroll
      for (int k = 0; k < NCH; ++k) {
        a[k] = mm > 0 ? ACOR(c0 + k) : 0.0;
        z[k] = ZN(0, c0 + k);
        w[k] = EWT(c0 + k);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < NCH; ++k) {
        const int i = c0 + k;
        const double bi = y[i];
        a[k] += bi;
        y[i] = z[k] + a[k];
        const double t = bi * w[k], u = a[k] * w[k];
        s2 += t * t;
        sa += u * u;
      }
#pragma unroll
      for (int k = 0; k < NCH; ++k) ACOR(c0 + k) = a[k];
    }
    const double del = sqrt(s2 / N);
    if (mm > 0) { const double r = del / m.delp; m.crate = (CRDOWN * m.crate > r) ? CRDOWN * m.crate : r; }
    const double dcon = del * (m.crate < 1.0 ? m.crate : 1.0) / m.tq4;
    if (dcon <= 1.0) {
      m.acnrm = (mm == 0) ? del : sqrt(sa / N);
      m.jcur = 0;
      return NLS_OK;
    }
    m.mnewt = mm + 1;
    if ((mm + 1 == MAXCOR) || ((mm + 1 >= 2) && (del > RDIV * m.delp)))
      return m.jcur ? NLS_CONV_FAIL : NLS_TRY_AGAIN;
    m.delp = del;
    return NLS_ITER;
  }

  // A successful step (CVODE's cvCompleteStep + cvPrepareNextStep), with one sweep over
  // the components for: zn[j] += l[j] acor; the saved correction zn[QMAX] = acor when
  // qwait reaches 1; the norms of the order-selection test (zn[q] and acor - cquot
  // zn[QMAX], weighted with this step's ewt); the error weights of the next step.
  // ---------------- array sweeps split across the wavefront: the four lanes of group g =
  // 0..3 of a cell each take NQ consecutive components.  Per component the arithmetic is
  // that of CVODE's cvCompleteStep / CVodeGetDky / the composed map; a sum over components (the
  // order-selection norms) is left as per-component terms in LDS and added in component
  // order by the integrator lane, so the results do not depend on the split.
  static constexpr int NGRP = 4, NQ = N / NGRP;
  static_assert(N % NGRP == 0, "whole groups");
  CEP_D BElem gblk(Rsrc r, int g, int kc) const
  {
    return BElem{r, (ln & 63u) * 8u + (unsigned)g * (NQ * CPB * 8u), kc * CPB * 8};
  }

  // scalar part of the completed step (cvCompleteStep) before its sweep (what the sweep and complete_end need
  // goes to the cell's LDS scalars)
  CEP_D void complete_begin(double dsm)
  {
    m.nst++;
    const int q = m.q;
#pragma unroll
    for (int i = QMAX; i >= 2; --i) if (i <= q) m.tau[i] = m.tau[i - 1];
    if ((q == 1) && (m.nst > 1)) m.tau[2] = m.tau[1];
    m.tau[1] = m.h;
    const bool save = (m.qwait == 2) && (q != B.maxord);   // qwait becomes 1
    m.qwait--;
    const bool norms = (m.etamax != 1.0) && (m.qwait == 0);
    const bool up = norms && (q != B.maxord) && (m.saved_tq5 != 0.0);
    m.c_save = save;
    m.c_dn = norms && q > 1;
    m.c_up = up;
    m.c_cquot = up ? (m.tq5 / m.saved_tq5) * hpow(m.h / m.tau[2], (double)(q + 1)) : 0.0;
    m.c_dsm = dsm;
  }
  // the completed step's sweep over group g's components; the norm terms go to t2 / t3
  template <typename TV>
  CEP_D void complete_part(int g, const TV &t2, const TV &t3) const
  {
    const int q = m.q;
    double l[QMAX + 1], oh[QMAX + 1];
#pragma unroll
    for (int j = 0; j <= QMAX; ++j) { l[j] = (j <= q) ? m.l[j] : 0.0; oh[j] = (j == q) ? 1.0 : 0.0; }
    const bool save = m.c_save;
    const double cquot = m.c_cquot, rt = B.rtol, at_ = B.atol;
    // all of the lane's components in one chunk: one memory round trip per sweep (chunks of
    // four components took four, 2.6% of the population's beat time)
    constexpr int C4 = NQ;
#pragma unroll
    for (int c0 = 0; c0 < NQ; c0 += C4) {
      constexpr int dummy = 0;
      (void)dummy;
      double z[C4][QMAX + 1], a[C4], w[C4];
#pragma unroll
      for (int k = 0; k < C4; ++k) {
        if (c0 + k < NQ) {
#pragma unroll
          for (int j = 0; j <= QMAX; ++j) z[k][j] = gblk(V.zn, g, j * N + c0 + k);
          a[k] = gblk(V.acor, g, c0 + k);
          w[k] = gblk(V.ewt, g, c0 + k);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < C4; ++k) {
        if (c0 + k < NQ) {
          const int i = g * NQ + c0 + k;
          const double zqm = z[k][QMAX];
          double zq = 0.0;
#pragma unroll
          for (int j = 0; j <= QMAX; ++j) {
            z[k][j] += l[j] * a[k];
            zq += oh[j] * z[k][j];
          }
          t2[i] = zq * w[k];
          t3[i] = (a[k] - cquot * zqm) * w[k];
          if (save) z[k][QMAX] = a[k];
          w[k] = 1.0 / (rt * fabs(z[k][0]) + at_);
        }
      }
#pragma unroll
      for (int k = 0; k < C4; ++k) {
        if (c0 + k < NQ) {
#pragma unroll
          for (int j = 0; j <= QMAX; ++j) gblk(V.zn, g, j * N + c0 + k) = z[k][j];
          gblk(V.ewt, g, c0 + k) = w[k];
        }
      }
    }
  }
  // the completed step after its sweep: the norms from the per-component terms, in component order
  template <typename TV>
  CEP_D void complete_end(const TV &t2, const TV &t3)
  {
    double s2 = 0.0, s3 = 0.0;
    // the terms are read in chunks (LDS reads issued together), summed in component order
    if (m.c_dn) {
#pragma unroll
      for (int c0 = 0; c0 < N; c0 += CH) {
        double t[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) t[k] = t2[c0 + k];
#pragma unroll
        for (int k = 0; k < CH; ++k) s2 += t[k] * t[k];
      }
    }
    if (m.c_up) {
#pragma unroll
      for (int c0 = 0; c0 < N; c0 += CH) {
        double t[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) t[k] = t3[c0 + k];
#pragma unroll
        for (int k = 0; k < CH; ++k) s3 += t[k] * t[k];
      }
    }
    if (m.c_save) m.saved_tq5 = m.tq5;
    prepare_next(m.c_dsm, s2, s3, m.c_dn, m.c_up);
    m.etamax = (m.nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  }
  // interpolate() (at s = m.i_s) and/or the attempt map (apply_map with M, entry k at Mp[k] in LDS) over group
  // g's components; with `ycopy` the interpolated values also become the cell's STATES
  template <typename YV>
  CEP_D void output_attempt_part(int g, bool interp, bool ycopy, const YV &yo, bool attempt, const YV &Mp,
                                 const YV &yp) const
  {
    const double s = m.i_s;
    const bool q5 = m.i_q5 != 0;
    double M[NMAP];
    if (attempt) {
#pragma unroll
      for (int k = 0; k < NMAP; ++k) M[k] = Mp[k];
    }
    constexpr int C4 = NQ;   // one chunk, one memory round trip (as in complete_part)
#pragma unroll
    for (int c0 = 0; c0 < NQ; c0 += C4) {
      double z[C4][QMAX + 1];
#pragma unroll
      for (int k = 0; k < C4; ++k)
        if (c0 + k < NQ) {
#pragma unroll
          for (int j = 0; j <= QMAX; ++j) z[k][j] = gblk(V.zn, g, j * N + c0 + k);
        }
      __builtin_amdgcn_sched_barrier(0);
      if (interp) {
#pragma unroll
        for (int k = 0; k < C4; ++k)
          if (c0 + k < NQ) {
            const int i = g * NQ + c0 + k;
            double v = q5 ? z[k][QMAX] : 0.0;
#pragma unroll
            for (int j = QMAX - 1; j >= 0; --j) v = z[k][j] + s * v;   // rows above q are zero
            yo[i] = v;
            if (ycopy) Y(i) = v;
          }
      }
      if (attempt) {
#pragma unroll
        for (int k = 0; k < C4; ++k)
          if (c0 + k < NQ) {
            const int i = g * NQ + c0 + k;
            double w[QMAX + 1];
#pragma unroll
            for (int r = 0; r <= QMAX; ++r) {
              double acc = 0.0;
#pragma unroll
              for (int col = QMAX; col >= r; --col) acc += M[map_idx(r, col)] * z[k][col];
              w[r] = acc;
            }
#pragma unroll
            for (int j = 0; j <= QMAX; ++j) gblk(V.zn, g, j * N + c0 + k) = w[j];
            yp[i] = w[0];
          }
      }
    }
  }

  CEP_D void set_eta()
  {
    if (m.eta < THRESH) {
      m.eta = 1.0;
      m.hprime = m.h;
    } else {
      if (m.eta > m.etamax) m.eta = m.etamax;
      const double r = fabs(m.h) * (1.0 / B.dt) * m.eta;
      m.eta /= (r > 1.0 ? r : 1.0);
      m.hprime = m.h * m.eta;
    }
  }

  CEP_D void prepare_next(double dsm, double s2, double s3, bool dn, bool up)
  {
    if (m.etamax == 1.0) {
      if (m.qwait < 2) m.qwait = 2;
      m.qprime = m.q;
      m.hprime = m.h;
      m.eta = 1.0;
      return;
    }
    const int q = m.q;
    const double etaq = 1.0 / (hpow(BIAS2 * dsm, 1.0 / (q + 1)) + ADDON);
    if (m.qwait != 0) {
      m.eta = etaq;
      m.qprime = q;
      set_eta();
      return;
    }
    m.qwait = 2;
    double etaqm1 = 0.0, etaqp1 = 0.0;
    if (dn) {
      const double ddn = sqrt(s2 / N) * m.tq1;
      etaqm1 = 1.0 / (hpow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    if (up) {
      const double dup = sqrt(s3 / N) * m.tq3;
      etaqp1 = 1.0 / (hpow(BIAS3 * dup, 1.0 / (q + 2)) + ADDON);
    }
    double etam = etaq;
    if (etaqm1 > etam) etam = etaqm1;
    if (etaqp1 > etam) etam = etaqp1;
    if (etam < THRESH) {
      m.eta = 1.0; m.qprime = q;
    } else if (etam == etaq) {
      m.eta = etaq; m.qprime = q;
    } else if (etam == etaqm1) {
      m.eta = etaqm1; m.qprime = q - 1;
    } else {
      m.eta = etaqp1; m.qprime = q + 1;
      double a[N];
#pragma unroll
      for (int i = 0; i < N; ++i) a[i] = ACOR(i);
#pragma unroll
      for (int i = 0; i < N; ++i) ZN(QMAX, i) = a[i];
    }
    set_eta();
  }

  // ---- initial step size (CVODE's cvHin): bounds, then up to MAX_ITERS trial RHS
  // evaluations at tn + hg.  Returns true when h is known without a trial.
  CEP_D bool hin_begin()
  {
    const double tdist = fabs(m.tout - m.tn);
    const double tround = UROUND * (fabs(m.tn) > fabs(m.tout) ? fabs(m.tn) : fabs(m.tout));
    m.hlb = HLB_FACTOR * tround;
    double hub_inv = 0.0;
#pragma unroll 4
    for (int i = 0; i < N; ++i) {
      const double tol = HUB_FACTOR * fabs(ZN(0, i)) + B.atol;
      const double r = fabs(ZN(1, i)) / tol;
      if (r > hub_inv) hub_inv = r;
    }
    double hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    m.hub = hub;
    m.hg = sqrt(m.hlb * hub);
    m.hin_count = 1;
    m.hin_ok = 0;
    if (hub < m.hlb) { hin_set(m.hg); return true; }
    return false;
  }
  // consume f(tn + hg, zn0 + hg zn1); returns true when the estimate is final
  template <typename FV>
  CEP_D bool hin_step(const FV &f)
  {
    const double hg = m.hg;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double t = (f[i] - ZN(1, i)) * (1.0 / hg) * EWT(i);
      s += t * t;
    }
    const double yddnrm = sqrt(s / N);
    double hnew;
    if (m.hin_ok || m.hin_count == MAX_ITERS) {
      hnew = hg;
    } else {
      hnew = (yddnrm * m.hub * m.hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * m.hub);
      const double hrat = hnew / hg;
      if ((hrat > 0.5) && (hrat < 2.0)) m.hin_ok = 1;
      if ((m.hin_count > 1) && (hrat > 2.0)) { hnew = hg; m.hin_ok = 1; }
      m.hg = hnew;
      m.hin_count++;
      return false;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < m.hlb) h0 = m.hlb;
    if (h0 > m.hub) h0 = m.hub;
    hin_set(h0);
    return true;
  }
  CEP_D void hin_set(double h)
  {
    const double rh = fabs(h) / B.dt;    // max step = output dt
    if (rh > 1.0) h /= rh;
    m.h = h; m.hscale = h; m.hprime = h;
    double z[N];
#pragma unroll
    for (int i = 0; i < N; ++i) z[i] = ZN(1, i) * h;
#pragma unroll
    for (int i = 0; i < N; ++i) ZN(1, i) = z[i];
  }
};

}  // namespace ordbdf
}  // namespace cep
