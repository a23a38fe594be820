// cep_gpu.hip -- kernels and C ABI of the cardiac cell-population engine (gfx950).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cep_cellsim.h"
#include "../../include/cep_tissue.h"
#include "adams.hpp"
#include "bdf_ord.hpp"
#include "driver_rules.hpp"
#include "ord11_model.hpp"
#include "ord11_host.hpp"
#include "tt06_model.hpp"

namespace cep {
namespace ord {
__constant__ MarkovTable g_markov;
}

using ordbdf::Buffers;

// ============================================================================ kernels

struct PrmAcc {
  const double *p;
  int pitch;
  CEP_D double operator()(int k) const { return p[k * pitch]; }
};

// computeRates for a batch: AoS states in, AoS rates / algebraic slots out
constexpr int ORD_NALG = 200;
template <typename R>
__global__ void __launch_bounds__(64) ord_rates_kernel(int n, int pitch, double t, const double *__restrict__ yin,
                                 double *__restrict__ fout, double *__restrict__ alg, const double *__restrict__ prm)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double y[ord::NS], f[ord::NS];
#pragma unroll
  for (int i = 0; i < ord::NS; ++i) y[i] = yin[(size_t)c * ord::NS + i];
  ord::Currents cur;
  const PrmAcc P{prm + c, pitch};
  ord::rhs<R>(t, y, f, P, &cur);
#pragma unroll
  for (int i = 0; i < ord::NS; ++i) fout[(size_t)c * ord::NS + i] = f[i];
  if (alg) {
    double *a = alg + (size_t)c * ORD_NALG;
    a[0] = cur.Istim; a[58] = cur.INa; a[60] = cur.INaL; a[66] = cur.Ito; a[83] = cur.ICaL;
    a[84] = cur.ICaNa; a[87] = cur.ICaK; a[90] = cur.IKr; a[96] = cur.IKs; a[98] = cur.IK1;
    a[130] = cur.INaCa_i; a[160] = cur.INaCa_ss; a[179] = cur.INaK; a[181] = cur.IKb;
    a[185] = cur.INab; a[189] = cur.ICab; a[190] = cur.IpCa; a[193] = cur.Jrel; a[198] = cur.Jup;
    a[42] = cur.CaMKa;
  }
}

constexpr int TT_NALG = 70;
template <typename R>
__global__ void __launch_bounds__(64) tt_rates_kernel(int n, int pitch, double t, const double *__restrict__ yin,
                                double *__restrict__ fout, double *__restrict__ alg, const double *__restrict__ prm)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double y[tt::NS], f[tt::NS];
#pragma unroll
  for (int i = 0; i < tt::NS; ++i) y[i] = yin[(size_t)c * tt::NS + i];
  tt::Currents cur;
  const PrmAcc P{prm + c, pitch};
  tt::rhs<R>(t, y, f, P, &cur);
#pragma unroll
  for (int i = 0; i < tt::NS; ++i) fout[(size_t)c * tt::NS + i] = f[i];
  if (alg) {
    double *a = alg + (size_t)c * TT_NALG;
    a[47] = cur.IK1; a[54] = cur.Ito; a[48] = cur.IKr; a[49] = cur.IKs; a[52] = cur.ICaL;
    a[55] = cur.INaK; a[50] = cur.INa; a[51] = cur.IbNa; a[56] = cur.INaCa; a[53] = cur.IbCa;
    a[58] = cur.IpK; a[57] = cur.IpCa; a[12] = cur.Istim;
  }
}

struct RunCfg {
  double dt;
  int nbeats_total, nwin, print_freq;
  double t_first;             // first output time (ORd-CiPA; dt unless set)
  double v_apd90, v_apd50;    // thresholds of the driver rules
  double *dr;                 // driver-rule state (DR_SN rows), or null: rules off
};

// win_cell: the cell's own first window beat (set when the cell was restarted on its own), or
// -1 for the batch's window, the last nwin of nbeats_total beats
template <class BM>
CEP_D void bm_config(BM &bm, double bcl, double stim_start, const RunCfg &rc, int cap, int win_cell)
{
  bm.bcl = bcl;
  bm.dt = rc.dt;
  bm.stim_start = stim_start;
  bm.beat_steps = (int)(bcl / rc.dt);
  const int nwin = rc.nwin > rc.nbeats_total ? rc.nbeats_total : rc.nwin;
  bm.win_start = win_cell >= 0 ? win_cell : rc.nbeats_total - nwin;
  bm.print_freq = rc.print_freq;
  bm.cap = cap;
}

// ---- ORd-CiPA: persistent CVODE-style BDF + biomarkers, CPW cells per wavefront.
//
// Lane roles inside the 64-lane workgroup (one wavefront):
//   lanes [0, CPW)        integrator of cell slot s = lane (BDF step logic, Newton)
//   lanes [CPW, 2*CPW)    output lane of slot s = lane - CPW: evaluates the right-hand side
//                         at each output point (CVode's y(tout) + computeRates) and runs
//                         the biomarker state machine
//   every lane            takes finite-difference Jacobian columns from a wave-level task
//                         queue whenever it has nothing else to evaluate in a pass
// Every pass runs the bookkeeping blocks (each masked to the lanes in that phase), then
// exactly one right-hand-side evaluation per busy lane, then consumes the results.  A
// cell's arithmetic never depends on its neighbours: results are the same alone and in
// any batch.
enum { RQ_NONE = 0, RQ_STEP, RQ_OUT, RQ_JAC };

CEP_D int lanes_below(unsigned long long mask)
{
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
}

// The array work of a pass split across the wavefront: lane l works on cell slot l % CPW,
// components [g*NQ, (g+1)*NQ) with g = l / CPW.  The first call does the completion sweeps
// (norm terms into the cell's staging vector and the output lane's), the second the
// interpolations to output points and the maps of the next attempts.
template <typename R, int CPW>
CEP_D void wave_sweeps(const ordbdf::Buffers &B, const ordbdf::BlockBufs &V, ordbdf::Mach *s_m, double (*s_v)[64],
                       int cell0, int lane, int pp, bool complete)
{
  using namespace ordbdf;
  static_assert(CPW * Lane<R>::NGRP == 64, "four lanes per cell");
  const int s = lane % CPW, g = lane / CPW;
  const int fl = (cell0 + s < B.ncell) ? s_m[s].wsw : 0;
  const LVec yv{&s_v[0][s]}, ov{&s_v[0][CPW + s]};
  if (complete) {
    if (fl & SWF_COMPLETE) {
      Lane<R> Ls{B, s_m[s], V, (unsigned)s, (unsigned)(cell0 + s), pp};
      Ls.complete_part(g, yv, ov);
    }
  } else if (fl & (SWF_INTERP | SWF_ATTEMPT)) {
    Lane<R> Ls{B, s_m[s], V, (unsigned)s, (unsigned)(cell0 + s), pp};
    Ls.output_attempt_part(g, (fl & SWF_INTERP) != 0, (fl & SWF_YCOPY) != 0, ov, (fl & SWF_ATTEMPT) != 0,
                           LVec{&s_v[0][2 * CPW + s]}, yv);
  }
}

// What a launch advances: LAUNCH_BEATS `n` beats of each cell's BCL, LAUNCH_OUTPUTS `n` further
// output points, LAUNCH_TO_OUTPUT up to the absolute output point `n` (cells already there take
// no step).  A template parameter rather than a run-time argument: the beat launch, the headline
// one, keeps the code of a kernel that reads no launch argument besides n (run-time forms of this
// choice gave FP64 builds whose Newton iterations failed from the first step; the three forms
// are compared bitwise by tests/test_adapter.py).
// LAUNCH_BEATS_ORDERED: whole beats, with workgroup w running cell group B.gorder[w] (groups
// dispatched longest beat first: the hardware dispatcher hands the next group to a CU as one
// finishes, so the long groups no longer trail at the end of a launch of more groups than fit
// at once)
enum { LAUNCH_BEATS = 0, LAUNCH_OUTPUTS = 1, LAUNCH_TO_OUTPUT = 2, LAUNCH_BEATS_ORDERED = 3 };

template <typename R, int CPW, int MODE>
__global__ void __launch_bounds__(64) ord_bdf_kernel(const Buffers B, const RunCfg rc, const int nbeats,
                                                     const long long max_passes)
{
  using namespace ordbdf;
  static_assert(2 * CPW <= 64, "an integrator and an output lane per cell");
  static_assert(CPW == CPB, "one block of integrator vectors per workgroup");
  // LDS is kept within 40 KiB, so that four one-wave workgroups share a CU (one per SIMD)
  __shared__ Mach s_m[CPW];
  __shared__ BiomarkerT<true> s_bm[CPW];          // the selected beat's record goes to HBM
  // each lane's right-hand-side argument, then its result: [component][lane].  The RHS
  // reads the argument into registers before it writes the result.  Between a pass's
  // consume and the next pass's Jacobian columns nothing else uses the columns of lanes
  // [2*CPW, 3*CPW): they hold the attempt maps (NMAP entries of cell slot s in column 2*CPW+s),
  // written by the integrators in the second bookkeeping stage and read by the second sweep.
  __shared__ double s_v[N][64];
  __shared__ double s_tout[CPW];
  __shared__ int s_oiout[CPW], s_opend[CPW], s_ofail[CPW];
  // Jacobian tasks: a cell publishes its NJC columns as one run, so task t is column t % NJC
  // of the cell slot in ring entry (t / NJC) % CPW (at most one run per cell is pending)
  __shared__ unsigned char s_jcell[CPW];
  // driver rules of the output lanes, then the output point's STATES[qnet, INaL_AUC, ICaL_AUC]:
  // dynamic LDS that only the launches with driver rules (rc.dr set) allocate
  extern __shared__ double s_dyn[];
  DriverRules *const s_dr = reinterpret_cast<DriverRules *>(s_dyn);
  double (*const s_acc)[3] = reinterpret_cast<double (*)[3]>(s_dyn + CPW * (sizeof(DriverRules) / sizeof(double)));
  static_assert(sizeof(DriverRules) % sizeof(double) == 0, "the STATES triples follow the rules");
  static_assert(3 * CPW <= 64 && NMAP <= N, "attempt maps in the staging columns of lanes [2*CPW, 3*CPW)");

  constexpr bool kBeats = MODE == LAUNCH_BEATS || MODE == LAUNCH_BEATS_ORDERED;
  const int lane = threadIdx.x;
  const int grp = MODE == LAUNCH_BEATS_ORDERED ? B.gorder[blockIdx.x] : (int)blockIdx.x;
  const int cell0 = grp * CPW;
  const bool is_int = lane < CPW && cell0 + lane < B.ncell;
  const bool is_out = lane >= CPW && lane < 2 * CPW && cell0 + lane - CPW < B.ncell;
  const int slot = lane < CPW ? lane : (lane < 2 * CPW ? lane - CPW : 0);
  const int c = cell0 + slot;
  const LVec LV{&s_v[0][lane]};                    // this lane's staging vector
  const double srur = sqrt(B.uround_fd);
  const size_t blk0 = (size_t)grp * CPB;            // first cell of this workgroup's block
  const BlockBufs V0{make_rsrc(B.zn + blk0 * ((QMAX + 1) * N), CPB * (QMAX + 1) * N * 8),
                     make_rsrc(B.ewt + blk0 * N, CPB * N * 8), make_rsrc(B.acor + blk0 * N, CPB * N * 8),
                     make_rsrc(B.ft + blk0 * N, CPB * N * 8), make_rsrc(B.jd + blk0 * (NJC * N), CPB * NJC * N * 8),
                     make_rsrc(B.fac + blk0 * NFAC, CPB * NFAC * 8)};

  // ---- launch prologue
  int trace_start = 0;
  if (is_int) {
    Lane<R> L{B, s_m[slot], V0, (unsigned)slot, (unsigned)c, B.pitch};
    Mach &m = s_m[slot];
    L.load();
    int ph = m.phase;
    if (m.status != ST_OK) {
      ph = PH_DONE;
    } else if (ph == PH_FRESH) {
      // CVodeInit at t0 = 0 with the current states
      for (int i = 0; i < N; ++i) L.ZN(0, i) = L.Y(i);
      m.tn = 0.0; m.tout = rc.t_first; m.q = 1; m.qprime = 1; m.qwait = 2; m.etamax = ETAMX1;
      m.h = 0.0; m.hprime = 0.0; m.hscale = 0.0; m.eta = 1.0; m.nst = m.nstlp = m.nstlj = 0;
      m.crate = 1.0; m.gamrat = 1.0; m.gammap = 0.0; m.saved_tq5 = 0.0; m.jcur = 0; m.iout = 0;
      m.nfe = m.nje = m.nsetups = m.nni = m.ncfn = m.netf = 0;
      for (int i = 0; i < 7; ++i) m.tau[i] = 0.0;
      m.tq1 = m.tq2 = m.tq3 = m.tq4 = m.tq5 = 0.0;
      m.ncf = m.nef = 0; m.nflag = FIRST_CALL; m.callSetup = 0; m.convfail = NO_FAILURES;
      ph = PH_INIT_F;
    } else {
      ph = PH_STEP_DONE;
    }
    m.phase = ph;
    m.nstloc = 0;
    m.jend = 0;
    m.sw_ops = 0;
    m.sw_dq = 0;
    m.sw_q1 = m.q;
    const int beat_steps = (int)(B.prm[ord::OP_BCL * B.pitch + c] / rc.dt);
    if constexpr (kBeats) {
      m.iout_end = m.iout + nbeats * beat_steps;
    } else {
      m.iout_end = MODE == LAUNCH_OUTPUTS ? m.iout + nbeats : nbeats;
      if (m.iout >= m.iout_end && ph != PH_INIT_F) { ph = PH_DONE; m.phase = ph; }
    }
  }
  // the output lane's rows of the selected beat's record
  double *const best_row = B.bmd + c + (size_t)(BM_D_N + BM_RES_N) * B.pitch;
  if (is_out) {
    BiomarkerT<true> &bm = s_bm[slot];
    // a cell's own window (cells restarted one by one) in the beat launches, which are the
    // ones that run such cells; the output-point launches keep the batch's window and the
    // code they had before the per-cell window existed (their FP64 build integrated nothing
    // with this one load added: the code of this kernel has been sensitive to such changes)
    bm_config(bm, B.prm[ord::OP_BCL * B.pitch + c], B.prm[ord::OP_STIM_START * B.pitch + c], rc, B.cap,
              kBeats ? B.win[c] : -1);
    const int iout0 = B.si[11 * B.pitch + c];
    const bool fresh = B.si[4 * B.pitch + c] == PH_FRESH;
    if (fresh) bm.begin(B.y[c], B.y[9 * B.pitch + c], best_row, B.pitch);
    else bm.load(B.bmd, B.bmi, B.pitch, c);
    if (rc.dr) {
      if (fresh) s_dr[slot].begin(B.y[c]);
      else s_dr[slot].load(rc.dr, B.pitch, c);
    }
    if constexpr (kBeats) {
      trace_start = iout0 + (nbeats - 1) * bm.beat_steps;
    } else {                                       // the trace holds the launch's last beat_steps outputs
      trace_start = (MODE == LAUNCH_OUTPUTS ? iout0 + nbeats : nbeats) - bm.beat_steps;
    }
    s_opend[slot] = 0;
    s_ofail[slot] = 0;
  }
  __syncthreads();

  int jq_head = 0, jq_tail = 0;                   // wave-uniform task queue cursors
  for (long long pass = 0;; ++pass) {
    int pp = B.pitch;
    const BlockBufs &V = V0;
    asm volatile("" : "+s"(pp));
    int req = RQ_NONE;
    double treq = 0.0;
    int ph = PH_DONE;
    // ================= integrator bookkeeping (each block masked to its phase), in two
    // stages around the wave-wide sweep of the completed steps
    if (is_int) {
      Lane<R> L{B, s_m[slot], V, (unsigned)slot, (unsigned)c, pp};
      Mach &m = s_m[slot];
      ph = m.phase;
      m.wsw = 0;
      if (s_ofail[slot] && ph != PH_DONE) { m.status = ST_NONFINITE; ph = PH_DONE; }
      if (pass >= max_passes && ph != PH_DONE) { m.status = ST_MAX_STEPS; ph = PH_DONE; }
      if (ph == PH_WAIT_JAC && jq_head >= m.jend) {
        m.nje++;
        // the staging vector served Jacobian columns meanwhile: f(tn, zn[0]) back from FT
        double fz[N];
        for (int i = 0; i < N; ++i) fz[i] = L.FT(i);
        for (int i = 0; i < N; ++i) LV[i] = fz[i];
        ph = PH_FACTOR;
      }
      if (ph == PH_FACTOR) {
        const bool ok = L.factor();
        m.nsetups++;
        m.callSetup = 0;
        m.gamrat = m.crate = 1.0;
        m.gammap = m.gamma;
        m.nstlp = m.nst;
        ph = ok ? PH_NEWTON : PH_CONVFAIL;
      }
      if (ph == PH_NEWTON) {
        const int r = L.newton(LV);
        if (r == NLS_OK) ph = PH_ETEST;
        else if (r == NLS_ITER) ph = PH_ITER;
        else if (r == NLS_TRY_AGAIN) { m.callSetup = 1; m.convfail = FAIL_BAD_J; ph = PH_F0; }
        else ph = PH_CONVFAIL;
      }
      if (ph == PH_CONVFAIL) {
        m.ncfn++;
        m.tn = m.saved_t;                     // restore; the array part runs in PREDICT
        m.ncf++;
        m.etamax = 1.0;
        if (m.ncf == MAXNCF) {
          m.status = ST_CONV_FAIL; ph = PH_DONE;
        } else {
          m.eta = ETACF;
          m.nflag = PREV_CONV_FAIL;
          m.sw_ops = Lane<R>::SW_RESTORE | Lane<R>::SW_RESCALE;
          m.sw_dq = 0; m.sw_q1 = m.q;
          ph = PH_PREDICT;
        }
      }
      if (ph == PH_ETEST) {
        const double dsm = m.acnrm * m.tq2;
        if (dsm <= 1.0) {
          L.complete_begin(dsm);               // the array update: wave sweep below
          m.wsw |= SWF_COMPLETE;
          ph = PH_STEP_DONE;
        } else {
          m.nef++;
          m.netf++;
          m.nflag = PREV_ERR_FAIL;
          m.tn = m.saved_t;                   // restore; the array part runs in PREDICT
          m.sw_ops = Lane<R>::SW_RESTORE;
          m.sw_dq = 0; m.sw_q1 = m.q;
          if (m.nef == MAXNEF) {
            m.status = ST_ERR_FAIL; ph = PH_DONE;
          } else {
            m.etamax = 1.0;
            if (m.nef <= MXNEF1) {
              double eta = 1.0 / (Lane<R>::hpow(BIAS2 * dsm, 1.0 / (m.q + 1)) + ADDON);
              if (eta < ETAMIN) eta = ETAMIN;
              if (m.nef >= SMALL_NEF && eta > ETAMXF) eta = ETAMXF;
              m.eta = eta;
              m.sw_ops |= Lane<R>::SW_RESCALE;
              ph = PH_PREDICT;
            } else if (m.q > 1) {
              m.eta = ETAMIN;
              m.sw_dq = -1;                    // cvAdjustOrder(-1), then q = q - 1
              m.sw_q1 = m.q - 1;
              m.qwait = m.q;
              m.sw_ops |= Lane<R>::SW_RESCALE;
              ph = PH_PREDICT;
            } else {
              m.eta = ETAMIN;
              m.h *= m.eta;
              m.hscale = m.h;
              m.qwait = LONG_WAIT;
              m.sw_ops |= Lane<R>::SW_RESTART;   // restore only, then f(tn, zn[0]) -> zn[1]
              ph = PH_PREDICT;
            }
          }
        }
      }
    }
    __syncthreads();
    wave_sweeps<R, CPW>(B, V, s_m, s_v, cell0, lane, pp, true);
    __syncthreads();
    if (is_int) {
      Lane<R> L{B, s_m[slot], V, (unsigned)slot, (unsigned)c, pp};
      Mach &m = s_m[slot];
      if (m.wsw & SWF_COMPLETE) L.complete_end(LV, LVec{&s_v[0][CPW + slot]});
      if (ph == PH_STEP_DONE) {
        if ((m.tn - m.tout) * m.h >= 0.0) {
          if (!s_opend[slot]) {              // hand y(tout) to the output lane
            m.i_s = (m.tout - m.tn) / m.h;       // interpolated by the wave sweep below
            m.i_q5 = m.q == QMAX;
            m.wsw |= SWF_INTERP;
            s_tout[slot] = m.tout;
            m.iout++;
            if (m.iout >= m.iout_end) m.wsw |= SWF_YCOPY;   // CellML STATES at the end of the run
            s_oiout[slot] = m.iout;
            s_opend[slot] = 1;
            m.nstloc = 0;
            m.tout += rc.dt;
            if (m.iout >= m.iout_end) ph = PH_DONE;
            else if ((m.tn - m.tout) * m.h < 0.0) ph = PH_START;
          }
        } else {
          ph = PH_START;
        }
      }
      if (ph == PH_START) {
        if (m.nstloc >= B.max_steps) {
          m.status = ST_MAX_STEPS; ph = PH_DONE;
        } else {
          // (the error weights for this step were set when the last step completed)
          m.nstloc++;
          m.sw_ops = 0; m.sw_dq = 0; m.sw_q1 = m.q;
          if (m.nst > 0 && m.hprime != m.h) {
            if (m.qprime != m.q) {
              m.sw_dq = m.qprime - m.q;
              m.sw_q1 = m.qprime;
              m.qwait = m.qprime + 1;
            }
            m.sw_ops = Lane<R>::SW_RESCALE;
          }
          m.saved_t = m.tn;
          m.ncf = m.nef = 0;
          m.nflag = FIRST_CALL;
          ph = PH_PREDICT;
        }
      }
      if (ph == PH_PREDICT) {
        // one sweep: [restore] [order change] [rescale] predict, y = zn[0]
        const int q0 = m.q, q1 = m.sw_q1;
        const bool restart = (m.sw_ops & Lane<R>::SW_RESTART) != 0;
        double ll[QMAX + 1], a1 = 0.0;
#pragma unroll
        for (int j = 0; j <= QMAX; ++j) ll[j] = 0.0;
        int ops = m.sw_ops;
        if (m.sw_dq != 0) ops |= L.order_ops(m.sw_dq, ll, a1);   // at the old order and hscale
        m.q = q1;
        if (ops & Lane<R>::SW_RESCALE) { m.h = m.hscale * m.eta; m.hscale = m.h; }
        if (!restart) { ops |= Lane<R>::SW_PREDICT; m.tn += m.h; }
        {
          double M[QMAX + 1][QMAX + 1];
          Lane<R>::build_map(ops, q0, q1, ll, a1, m.eta, M);
#pragma unroll
          for (int r = 0; r <= QMAX; ++r)
#pragma unroll
            for (int col = r; col <= QMAX; ++col) s_v[map_idx(r, col)][2 * CPW + slot] = M[r][col];
          m.wsw |= SWF_ATTEMPT;                // applied by the wave sweep below
        }
        m.sw_ops = 0; m.sw_dq = 0; m.sw_q1 = m.q;
        if (restart) {
          ph = PH_RESTART_F;
        } else {
          L.set_bdf();
          m.callSetup = (m.nflag == PREV_CONV_FAIL) || (m.nflag == PREV_ERR_FAIL) || (m.nst == 0) ||
                        (m.nst >= m.nstlp + MSBP) || (fabs(m.gamrat - 1.0) > DGMAX);
          m.convfail = ((m.nflag == FIRST_CALL) || (m.nflag == PREV_ERR_FAIL)) ? NO_FAILURES : FAIL_OTHER;
          ph = PH_F0;
        }
        treq = m.tn;
        req = RQ_STEP;                         // y = zn[0] from the sweep
      } else if (ph == PH_F0 || ph == PH_INIT_F) {
        double z[N];
        for (int i = 0; i < N; ++i) z[i] = L.ZN(0, i);
        for (int i = 0; i < N; ++i) LV[i] = z[i];
        treq = m.tn;
        req = RQ_STEP;
      } else if (ph == PH_ITER) {
        treq = m.tn;                         // y was set by the Newton block
        req = RQ_STEP;
      } else if (ph == PH_HIN) {
        double z[N];
        for (int i = 0; i < N; ++i) z[i] = m.hg * L.ZN(1, i) + L.ZN(0, i);
        for (int i = 0; i < N; ++i) LV[i] = z[i];
        treq = m.tn + m.hg;
        req = RQ_STEP;
      }
      m.phase = ph;
    }
    __syncthreads();
    wave_sweeps<R, CPW>(B, V, s_m, s_v, cell0, lane, pp, false);
    __syncthreads();
    // ================= output lanes
    if (is_out && s_opend[slot]) {                 // y(tout) is already in s_v
      treq = s_tout[slot];
      req = RQ_OUT;
    }
    // ================= Jacobian columns for every other lane
    const unsigned long long freem = __ballot(req == RQ_NONE);
    const int navail = jq_tail - jq_head;
    int jown = 0, jcol = 0;
    double jinc = 0.0;
    if (req == RQ_NONE) {
      const int r = lanes_below(freem);
      if (r < navail) {
        const unsigned t = (unsigned)(jq_head + r);
        jown = s_jcell[(t / NJC) % CPW];
        jcol = (int)(t % NJC);
        const int co = cell0 + jown;
        Lane<R> Lo{B, s_m[jown], V, (unsigned)jown, (unsigned)co, pp};
        const int k = jstate(jcol);
        const double ysave = Lo.ZN_lane(0, (unsigned)k);
        double inc = srur * fabs(ysave);
        const double alt = s_m[jown].mininc / Lo.EWT_lane((unsigned)k);
        if (alt > inc) inc = alt;
        // no runtime index into the register array: select the perturbed component
        double z[N];
        for (int i = 0; i < N; ++i) z[i] = (i == k) ? ysave + inc : Lo.ZN(0, i);
        for (int i = 0; i < N; ++i) LV[i] = z[i];
        jinc = inc;
        treq = s_m[jown].tn;
        req = RQ_JAC;
      }
    }
    {
      const int nfree = __popcll(freem);
      jq_head += nfree < navail ? nfree : navail;
    }
    if (__ballot(req != RQ_NONE) == 0ull) break;      // every cell finished
    // ================= one right-hand side per busy lane
    ord::Currents cur;
    double v0 = 0.0, ca0 = 0.0;
    if (req != RQ_NONE) {
      const int cc = req == RQ_JAC ? cell0 + jown : c;
      const PrmAcc P{B.prm + cc, pp};
      double y[N];
      for (int i = 0; i < N; ++i) y[i] = LV[i];
      v0 = y[0];
      ca0 = y[9];
      if (req == RQ_OUT && rc.dr) {
        s_acc[slot][0] = y[49]; s_acc[slot][1] = y[50]; s_acc[slot][2] = y[51];
      }
      ord::rhs<R>(treq, y, LV, P, &cur);          // f overwrites the argument in LDS
    }
    const LVec &f = LV;
    // ================= consume
    bool publish = false;
    if (req == RQ_STEP) {
      Lane<R> L{B, s_m[slot], V, (unsigned)slot, (unsigned)c, pp};
      Mach &m = s_m[slot];
      m.nfe++;
      int ph = m.phase;
      bool fin = true;
      for (int i = 0; i < N; ++i) fin &= (bool)isfinite(f[i]);   // no short-circuit branches
      if (!fin) {
        m.status = ST_NONFINITE; ph = PH_DONE;
      } else if (ph == PH_F0) {
        m.mnewt = 0;
        if (m.callSetup) {
          const double dgamma = fabs((m.gamma / m.gammap) - 1.0);
          const bool jbad = (m.nst == 0) || (m.nst > m.nstlj + MSBJ) ||
                            ((m.convfail == FAIL_BAD_J) && (dgamma < DGMAX_J)) || (m.convfail == FAIL_OTHER);
          if (jbad) {
            m.nstlj = m.nst;
            m.jcur = 1;
            double w[N];
            for (int i = 0; i < N; ++i) w[i] = L.EWT(i);
            double s = 0.0;
            for (int i = 0; i < N; ++i) { const double t = f[i] * w[i]; s += t * t; }
            for (int i = 0; i < N; ++i) L.FT(i) = f[i];
            const double fnorm = sqrt(s / N);
            m.mininc = (fnorm != 0.0) ? (1000.0 * fabs(m.h) * B.uround_fd * N * fnorm) : 1.0;
            m.nfe += NJC;
            publish = true;
            ph = PH_WAIT_JAC;
          } else {
            m.jcur = 0;
            ph = PH_FACTOR;                    // f stays in the staging vector for NEWTON
          }
        } else {
          ph = PH_NEWTON;
        }
      } else if (ph == PH_ITER) {
        ph = PH_NEWTON;
      } else if (ph == PH_RESTART_F) {
        for (int i = 0; i < N; ++i) L.ZN(1, i) = m.h * f[i];
        ph = PH_PREDICT;
      } else if (ph == PH_INIT_F) {
        for (int i = 0; i < N; ++i) L.ZN(1, i) = f[i];
        L.set_ewt();
        ph = L.hin_begin() ? PH_START : PH_HIN;
      } else if (ph == PH_HIN) {
        ph = L.hin_step(f) ? PH_START : PH_HIN;
      }
      m.phase = ph;
    } else if (req == RQ_OUT) {
      // CVode returned y(tout); computeRates at tout feeds the biomarkers
      bool fin = true;
      for (int i = 0; i < N; ++i) fin &= (bool)isfinite(f[i]);   // no short-circuit branches
      if (!fin) s_ofail[slot] = 1;
      const int iout = s_oiout[slot];
      TsView tv = B.ts;
      if (tv.t) { tv.t += c; tv.vm += c; tv.ca += c; tv.ires += c; tv.rp += c; tv.rv += c; }
      const double ires[6] = { cur.INaL, cur.ICaL, cur.Ito, cur.IKr, cur.IKs, cur.IK1 };
      s_bm[slot].sample(iout, treq, v0, ca0, f[0],
                        cur.INaL + cur.ICaL + cur.Ito + cur.IKr + cur.IKs + cur.IK1, cur.INaL, cur.ICaL,
                        B.ct + c, B.cv + c, pp, tv, ires, best_row);
      if (B.vtrace && iout > trace_start && iout - trace_start <= B.tcap)
        B.vtrace[(iout - trace_start - 1) * pp + c] = v0;
      if (rc.dr) {
        // benchpatch takes RATES[V] at the next output time: the same currents, the stimulus there
        const PrmAcc P{B.prm + c, pp};
        const double bcl = P(ord::OP_BCL);
        auto istim = [&](double tt) {
          return ord::stimulus(tt, P(ord::OP_STIM_START), P(ord::OP_STIM_END), P(ord::OP_STIM_AMP), bcl,
                               P(ord::OP_STIM_DUR));
        };
        s_dr[slot].sample(iout, treq, v0, ca0, s_acc[slot][0], s_acc[slot][1], s_acc[slot][2],
                          f[0] + istim(treq) - istim(treq + rc.dt), bcl, rc.dt, rc.nbeats_total, rc.v_apd90,
                          rc.v_apd50);
      }
      s_opend[slot] = 0;
    } else if (req == RQ_JAC) {
      Lane<R> Lo{B, s_m[jown], V, (unsigned)jown, (unsigned)(cell0 + jown), pp};
      const double iinv = 1.0 / jinc;
      double ft[N], fv[N];
      for (int i = 0; i < N; ++i) ft[i] = Lo.FT(i);
      for (int i = 0; i < N; ++i) fv[i] = f[i];
      // only the entries the factorisation reads reach memory: the others get an offset past the
      // block's buffer range, where the hardware's range check drops the store (no branches)
      const int kst = jstate(jcol);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        BElem e = Lo.JD_lane((unsigned)jcol, i);
        e.voff += jac_read(i, kst) ? 0u : 0x40000000u;
        e = iinv * fv[i] - iinv * ft[i];
      }
    }
    // integrators that need a fresh Jacobian queue their 48 columns
    const unsigned long long pubm = __ballot(publish);
    if (publish) {
      const int run = jq_tail / NJC + lanes_below(pubm);   // jq_tail is a multiple of NJC
      s_jcell[run % CPW] = (unsigned char)slot;
      s_m[slot].jend = (run + 1) * NJC;
    }
    jq_tail += NJC * __popcll(pubm);
    __syncthreads();
  }
  // ---- launch epilogue
  if (is_int) {
    Lane<R> L{B, s_m[slot], V0, (unsigned)slot, (unsigned)c, B.pitch};
    Mach &m = s_m[slot];
    if (m.phase == PH_DONE && m.status == ST_OK) m.phase = PH_STEP_DONE;
    L.store();
  }
  if (is_out) {
    s_bm[slot].store(B.bmd, B.bmi, B.pitch, c);
    if (rc.dr) s_dr[slot].store(rc.dr, B.pitch, c);
  }
}

// ---- TT06: fixed-dt Rush-Larsen + biomarkers, one cell per lane
struct TtBuffers {
  int ncell, pitch, cap, tcap;
  double *y, *prm, *bmd, *ct, *cv, *tcur, *vtrace;
  int *bmi, *iout, *status;
  const int *win;     // per-cell first window beat, -1: the batch's window
};

// WPE: waves per SIMD the register allocation leaves room for (the register split).  WPE 2: at
// most 256 registers, spills to scratch; 200,000 cells make 3125 waves, which with one wave per
// SIMD took four rounds of the 1024 SIMDs.  Measured per beat: FP32 rates 0.60 s against 0.78 s
// (320 registers, 220 bytes of spills per lane once capped), FP64 1.39 s against 1.64 s (390
// registers, 512 bytes of spills): the second wave hides more than the spills cost.  WPE 1: all
// 512 registers, no spills, for batches with no more waves than SIMDs.
template <typename R, int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
tt_rl_kernel(TtBuffers B, RunCfg rc, int nbeats)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B.ncell || B.status[c] == ST_IDLE) return;
  const PrmAcc P{B.prm + c, B.pitch};
  double y[tt::NS];
#pragma unroll
  for (int i = 0; i < tt::NS; ++i) y[i] = B.y[(size_t)i * B.pitch + c];
  Biomarker bm;
  bm_config(bm, P(tt::TP_BCL), P(tt::TP_STIM_START), rc, B.cap, B.win[c]);
  int iout = B.iout[c];
  double tcurr = B.tcur[c];
  if (iout == 0) bm.begin(y[0], y[3]); else bm.load(B.bmd, B.bmi, B.pitch, c);
  double *ct = B.ct + c, *cv = B.cv + c;
  const int steps = nbeats * bm.beat_steps;
  const int trace_start = iout + steps - bm.beat_steps;
  // population_tn2006's rules (rc.dr set): crossing APD and the last beat's mean Ca_i
  TnRules tn{y[0], 0.0, 0.0, 0.0, 0.0, 0.0};
  const double tn_last = (rc.nbeats_total - 1) * P(tt::TP_BCL);
  if (rc.dr && iout > 0) {
    double *t6 = &tn.v_prev;
    for (int k = 0; k < 6; ++k) t6[k] = rc.dr[(size_t)k * B.pitch + c];
  }
  for (int s = 0; s < steps; ++s) {
    tt::Currents cur;
    double dvdt;
    tt::rush_larsen<R>(tcurr, rc.dt, y, P, dvdt, &cur);
    tcurr += rc.dt;
    iout++;
    bm.sample(iout, tcurr, y[0], y[3], dvdt, cur.ICaL + cur.Ito + cur.IKr + cur.IKs + cur.IK1, 0.0,
              cur.ICaL, ct, cv, B.pitch);
    if (rc.dr) tn.sample(tcurr, y[0], y[3], tn_last, rc.v_apd90);
    if (B.vtrace && iout > trace_start && iout - trace_start <= B.tcap)
      B.vtrace[(size_t)(iout - trace_start - 1) * B.pitch + c] = y[0];
  }
  bool fin = true;
#pragma unroll
  for (int i = 0; i < tt::NS; ++i) { B.y[(size_t)i * B.pitch + c] = y[i]; fin = fin && isfinite(y[i]); }
  if (!fin) B.status[c] = ST_NONFINITE;
  B.iout[c] = iout;
  B.tcur[c] = tcurr;
  bm.store(B.bmd, B.bmi, B.pitch, c);
  if (rc.dr) {
    const double *t6 = &tn.v_prev;
    for (int k = 0; k < 6; ++k) rc.dr[(size_t)k * B.pitch + c] = t6[k];
  }
}

// ---- ORd 2011: the DrugSimulation loop (drug_sim.cpp:306-420) -- adaptive Rush-Larsen steps
// clipped to beat ends, per-beat features over the trailing nwin beats -- one cell per lane.
// Persistent per-cell scalars (biomarker doubles area): [0, O11_SD_N) the loop's running
// values, [O11_BEST, O11_BEST + 16) the selected beat's record; ints: pace, eligible AP,
// steepest pace, initialised flag.
enum : int {
  O11_T = 0, O11_R30, O11_R50, O11_R90, O11_TPEAK, O11_QAP, O11_Q4AP, O11_INALAP, O11_ICALAP,
  O11_QCL, O11_Q4CL, O11_INALCL, O11_ICALCL, O11_TMP_DVDT, O11_TMP_PEAK, O11_TMP_VALLEY, O11_SD_N,
  O11_BEST = O11_SD_N
};
enum : int { O11_I_PACE = 0, O11_I_ELIG, O11_I_STEEP, O11_I_INIT };
static_assert(O11_BEST + CEP_NBM <= BM_TOTAL_D, "ORd 2011 loop state fits the biomarker area");
static_assert(O11_I_INIT < BM_I_N, "ORd 2011 loop ints fit the biomarker area");

struct O11Buffers {
  int ncell, pitch, nbeats_total, nwin;
  double *y, *prm, *bmd;
  int *bmi, *status;
  long long *stats;
  double time_point, max_dt;
};

// WPE (the register split): 2 = at most 256 registers, two waves per SIMD (the large-population
// kernel: 100,000 cells make 1563 waves, more than the 1024 SIMDs).  Measured on 100,000 cells,
// FP64: 371 ms per beat against 423 ms with 451 registers and one wave, despite 776 bytes of
// spills per lane.  1 = all 512 registers and no spills.
template <typename R, int WPE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE)))
ord11_rl_kernel(O11Buffers B, int beats_end, long long max_steps)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B.ncell) return;
  if (B.status[c] != ST_OK) return;
  const size_t P_ = (size_t)B.pitch;
  const PrmAcc P{B.prm + c, B.pitch};
  double y[ord11::NS];
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) y[i] = B.y[i * P_ + c];
  double *sd = B.bmd + c;
  int *si = B.bmi + c;
  double st[O11_SD_N];
#pragma unroll
  for (int k = 0; k < O11_SD_N; ++k) st[k] = sd[k * P_];
  int pace = si[O11_I_PACE * P_], elig = si[O11_I_ELIG * P_], steep = si[O11_I_STEEP * P_];
  double best_dvdt = sd[(O11_BEST + CEP_BM11_DVMDT_REPOL) * P_];
  if (si[O11_I_INIT * P_] == 0) {
    // cipa_result.init / temp_result.init with the starting Vm (drug_sim.cpp:166-167)
    st[O11_TMP_DVDT] = -999.0; st[O11_TMP_PEAK] = -999.0; st[O11_TMP_VALLEY] = y[0];
    for (int k = 0; k < CEP_NBM; ++k) sd[(O11_BEST + k) * P_] = 0.0;
    sd[(O11_BEST + CEP_BM11_DVMDT_REPOL) * P_] = -999.0;
    sd[(O11_BEST + CEP_BM11_VM_PEAK) * P_] = -999.0;
    sd[(O11_BEST + CEP_BM11_VM_VALLEY) * P_] = y[0];
    best_dvdt = -999.0;
    si[O11_I_INIT * P_] = 1;
  }
  const double bcl = P(ord11::P_BCL);
  const double tmax = beats_end * bcl;
  const int first_win = B.nbeats_total - B.nwin;
  double t = st[O11_T];
  long long n = 0;
  while (t < tmax && n < max_steps) {
    double f[ord11::NS];
    ord11::Gates g;
    ord11::Currents cur;
    ord11::rhs<R>(t, y, f, P, cur, g);
    const double dts = ord11::time_step(t, B.time_point, B.max_dt, bcl, f[0]);
    double dt;
    if (floor((t + dts) / bcl) == floor(t / bcl)) {
      dt = dts;
    } else {
      dt = (floor(t / bcl) + 1) * bcl - t;
      if (elig && pace >= first_win) {
        if (st[O11_TMP_DVDT] > best_dvdt) {
          best_dvdt = st[O11_TMP_DVDT];
          steep = pace;
          double *bs = sd + O11_BEST * P_;
          bs[CEP_BM11_QNET_AP * P_] = st[O11_QAP];
          bs[CEP_BM11_QNET4_AP * P_] = st[O11_Q4AP];
          bs[CEP_BM11_INAL_AUC_AP * P_] = st[O11_INALAP];
          bs[CEP_BM11_ICAL_AUC_AP * P_] = st[O11_ICALAP];
          bs[CEP_BM11_QNET_CL * P_] = st[O11_QCL];
          bs[CEP_BM11_QNET4_CL * P_] = st[O11_Q4CL];
          bs[CEP_BM11_INAL_AUC_CL * P_] = st[O11_INALCL];
          bs[CEP_BM11_ICAL_AUC_CL * P_] = st[O11_ICALCL];
          bs[CEP_BM11_VM_PEAK * P_] = st[O11_TMP_PEAK];
          bs[CEP_BM11_VM_VALLEY * P_] = st[O11_TMP_VALLEY];
          bs[CEP_BM11_DVMDT_REPOL * P_] = st[O11_TMP_DVDT];
          bs[CEP_BM11_T_PEAK * P_] = st[O11_TPEAK];
          bs[CEP_BM11_VM_REPOL30 * P_] = st[O11_R30];
          bs[CEP_BM11_VM_REPOL50 * P_] = st[O11_R50];
          bs[CEP_BM11_VM_REPOL90 * P_] = st[O11_R90];
        }
      }
#pragma unroll
      for (int k = O11_QAP; k <= O11_ICALCL; ++k) st[k] = 0.0;
      st[O11_TPEAK] = 0.0;
      st[O11_TMP_DVDT] = -999.0; st[O11_TMP_PEAK] = -999.0; st[O11_TMP_VALLEY] = y[0];
      pace++;
      elig = 0;
    }
    ord11::rl_step<R>(dt, y, f, g);
    ++n;
    if (pace >= first_win) {
      const double t_beat = bcl * pace + ord11::stim_start;
      if (t > t_beat + 2 && t < t_beat + 10 && fabs(cur.INa) < 1) {
        if (y[0] > st[O11_TMP_PEAK]) {
          st[O11_TMP_PEAK] = y[0];
          if (y[0] > 0) {
            const double amp = st[O11_TMP_PEAK] - st[O11_TMP_VALLEY];
            st[O11_R30] = st[O11_TMP_PEAK] - 0.3 * amp;
            st[O11_R50] = st[O11_TMP_PEAK] - 0.5 * amp;
            st[O11_R90] = st[O11_TMP_PEAK] - 0.9 * amp;
            elig = 1;
            st[O11_TPEAK] = t;
          } else {
            elig = 0;
          }
        }
      } else if (t > t_beat + 10 && elig) {
        if (f[0] > st[O11_TMP_DVDT] && y[0] <= st[O11_R30] && y[0] >= st[O11_R90]) st[O11_TMP_DVDT] = f[0];
      }
      const double inet = cur.INaL + cur.ICaL + cur.Ito + cur.IKr + cur.IKs + cur.IK1;
      const double inet4 = cur.INaL + cur.ICaL + cur.IKr + cur.INa;
      if (elig && y[0] > st[O11_R90]) {
        st[O11_QAP] += (inet * dt) / 1000.;
        st[O11_Q4AP] += (inet4 * dt) / 1000.;
        st[O11_INALAP] += cur.INaL * dt;
        st[O11_ICALAP] += cur.ICaL * dt;
      }
      st[O11_QCL] += (inet * dt) / 1000.;
      st[O11_Q4CL] += (inet4 * dt) / 1000.;
      st[O11_INALCL] += cur.INaL * dt;
      st[O11_ICALCL] += cur.ICaL * dt;
    }
    t += dt;
  }
  st[O11_T] = t;
  bool fin = true;
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) { B.y[i * P_ + c] = y[i]; fin = fin && isfinite(y[i]); }
  if (!fin) B.status[c] = ST_NONFINITE;
  else if (t < tmax) B.status[c] = ST_MAX_STEPS;
#pragma unroll
  for (int k = 0; k < O11_SD_N; ++k) sd[k * P_] = st[k];
  si[O11_I_PACE * P_] = pace; si[O11_I_ELIG * P_] = elig; si[O11_I_STEEP * P_] = steep;
  B.stats[c] += n;              // steps (one right-hand side each)
}

// The loop's bookkeeping shared by the split kernels below (the one-wavefront kernel above has
// the same code inline): the beat-end selection of the steepest beat's features, and the
// per-step features (drug_sim.cpp:306-420).
struct O11Loop {
  double st[O11_SD_N];
  int pace, elig, steep;
  double best_dvdt;
};

CEP_D void o11_beat_end(O11Loop &L, double *sd, size_t P_, int first_win, double v0)
{
  double *st = L.st;
  if (L.elig && L.pace >= first_win) {
    if (st[O11_TMP_DVDT] > L.best_dvdt) {
      L.best_dvdt = st[O11_TMP_DVDT];
      L.steep = L.pace;
      double *bs = sd + O11_BEST * P_;
      bs[CEP_BM11_QNET_AP * P_] = st[O11_QAP];
      bs[CEP_BM11_QNET4_AP * P_] = st[O11_Q4AP];
      bs[CEP_BM11_INAL_AUC_AP * P_] = st[O11_INALAP];
      bs[CEP_BM11_ICAL_AUC_AP * P_] = st[O11_ICALAP];
      bs[CEP_BM11_QNET_CL * P_] = st[O11_QCL];
      bs[CEP_BM11_QNET4_CL * P_] = st[O11_Q4CL];
      bs[CEP_BM11_INAL_AUC_CL * P_] = st[O11_INALCL];
      bs[CEP_BM11_ICAL_AUC_CL * P_] = st[O11_ICALCL];
      bs[CEP_BM11_VM_PEAK * P_] = st[O11_TMP_PEAK];
      bs[CEP_BM11_VM_VALLEY * P_] = st[O11_TMP_VALLEY];
      bs[CEP_BM11_DVMDT_REPOL * P_] = st[O11_TMP_DVDT];
      bs[CEP_BM11_T_PEAK * P_] = st[O11_TPEAK];
      bs[CEP_BM11_VM_REPOL30 * P_] = st[O11_R30];
      bs[CEP_BM11_VM_REPOL50 * P_] = st[O11_R50];
      bs[CEP_BM11_VM_REPOL90 * P_] = st[O11_R90];
    }
  }
#pragma unroll
  for (int k = O11_QAP; k <= O11_ICALCL; ++k) st[k] = 0.0;
  st[O11_TPEAK] = 0.0;
  st[O11_TMP_DVDT] = -999.0; st[O11_TMP_PEAK] = -999.0; st[O11_TMP_VALLEY] = v0;
  L.pace++;
  L.elig = 0;
}

CEP_D void o11_features(O11Loop &L, int first_win, double bcl, double t, double dt, double v, double dvdt,
                        const ord11::Currents &cur)
{
  double *st = L.st;
  if (L.pace < first_win) return;
  const double t_beat = bcl * L.pace + ord11::stim_start;
  if (t > t_beat + 2 && t < t_beat + 10 && fabs(cur.INa) < 1) {
    if (v > st[O11_TMP_PEAK]) {
      st[O11_TMP_PEAK] = v;
      if (v > 0) {
        const double amp = st[O11_TMP_PEAK] - st[O11_TMP_VALLEY];
        st[O11_R30] = st[O11_TMP_PEAK] - 0.3 * amp;
        st[O11_R50] = st[O11_TMP_PEAK] - 0.5 * amp;
        st[O11_R90] = st[O11_TMP_PEAK] - 0.9 * amp;
        L.elig = 1;
        st[O11_TPEAK] = t;
      } else {
        L.elig = 0;
      }
    }
  } else if (t > t_beat + 10 && L.elig) {
    if (dvdt > st[O11_TMP_DVDT] && v <= st[O11_R30] && v >= st[O11_R90]) st[O11_TMP_DVDT] = dvdt;
  }
  const double inet = cur.INaL + cur.ICaL + cur.Ito + cur.IKr + cur.IKs + cur.IK1;
  const double inet4 = cur.INaL + cur.ICaL + cur.IKr + cur.INa;
  if (L.elig && v > st[O11_R90]) {
    st[O11_QAP] += (inet * dt) / 1000.;
    st[O11_Q4AP] += (inet4 * dt) / 1000.;
    st[O11_INALAP] += cur.INaL * dt;
    st[O11_ICALAP] += cur.ICaL * dt;
  }
  st[O11_QCL] += (inet * dt) / 1000.;
  st[O11_Q4CL] += (inet4 * dt) / 1000.;
  st[O11_INALCL] += cur.INaL * dt;
  st[O11_ICALCL] += cur.ICaL * dt;
}

// loads the loop state of cell c (and initialises it on the first launch); false if the
// cell does not run
CEP_D void o11_load_loop(O11Loop &L, const O11Buffers &B, int c, size_t P_, double v0)
{
  double *sd = B.bmd + c;
  int *si = B.bmi + c;
#pragma unroll
  for (int k = 0; k < O11_SD_N; ++k) L.st[k] = sd[k * P_];
  L.pace = si[O11_I_PACE * P_]; L.elig = si[O11_I_ELIG * P_]; L.steep = si[O11_I_STEEP * P_];
  L.best_dvdt = sd[(O11_BEST + CEP_BM11_DVMDT_REPOL) * P_];
  if (si[O11_I_INIT * P_] == 0) {
    L.st[O11_TMP_DVDT] = -999.0; L.st[O11_TMP_PEAK] = -999.0; L.st[O11_TMP_VALLEY] = v0;
    for (int k = 0; k < CEP_NBM; ++k) sd[(O11_BEST + k) * P_] = 0.0;
    sd[(O11_BEST + CEP_BM11_DVMDT_REPOL) * P_] = -999.0;
    sd[(O11_BEST + CEP_BM11_VM_PEAK) * P_] = -999.0;
    sd[(O11_BEST + CEP_BM11_VM_VALLEY) * P_] = v0;
    L.best_dvdt = -999.0;
    si[O11_I_INIT * P_] = 1;
  }
}

// stores the states, status, loop state and step count of cell c
CEP_D void o11_store(const O11Buffers &B, int c, size_t P_, const double *y, O11Loop &L, double t, double tmax,
                     long long n)
{
  double *sd = B.bmd + c;
  int *si = B.bmi + c;
  L.st[O11_T] = t;
  bool fin = true;
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) { B.y[i * P_ + c] = y[i]; fin = fin && isfinite(y[i]); }
  if (!fin) B.status[c] = ST_NONFINITE;
  else if (t < tmax) B.status[c] = ST_MAX_STEPS;
#pragma unroll
  for (int k = 0; k < O11_SD_N; ++k) sd[k * P_] = L.st[k];
  si[O11_I_PACE * P_] = L.pace; si[O11_I_ELIG * P_] = L.elig; si[O11_I_STEEP * P_] = L.steep;
  B.stats[c] += n;              // steps (one right-hand side each)
}

// The step of each cell split over the four wavefronts of a 256-thread workgroup (lane l of
// each holds cell blockIdx.x * 64 + l), on four SIMDs:
//   wave 0: the INa / INaL / Ito gate kinetics and their Rush-Larsen update (gates 10-24);
//   wave 1: the ICaL / IKr / IKs / IK1 gate kinetics and their update (gates 25-38);
//   wave 2: the Na and K currents;
//   wave 3: the Ca currents, pumps, exchangers and fluxes, then the rates, the step size, the
//           Jrel / Euler update of V and the concentrations, and the features.
// Per step, three barriers: after the rates and currents (wave 2's currents to LDS), after
// the step size (to LDS), after the updates (gates, V and concentrations to LDS).  Every wave
// sees the same t and step count for every lane, so all run the same number of passes.  Per
// cell the arithmetic is that of ord11_rl_kernel.
template <typename R>
__global__ void __launch_bounds__(256) ord11_rl4_kernel(O11Buffers B, int beats_end, long long max_steps)
{
  constexpr int NGT = 38 - ord11::G0 + 1;       // gates 10..38 (incl. nca)
  __shared__ double s_gate[NGT][64];
  __shared__ double s_core[10][64];             // V, CaMKt and the concentrations
  __shared__ double s_nak[8][64];               // wave 2's currents
  __shared__ double s_dt[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  const bool valid = c < B.ncell && B.status[c] == ST_OK;
  const size_t P_ = (size_t)B.pitch;
  const PrmAcc P{B.prm + (valid ? c : 0), B.pitch};
  double y[ord11::NS];
  O11Loop L;
  double t = 0.0, bcl = 1.0;
  if (valid) {
#pragma unroll
    for (int i = 0; i < ord11::NS; ++i) y[i] = B.y[i * P_ + c];
    bcl = P(ord11::P_BCL);
    t = B.bmd[c + O11_T * P_];
    if (wave == 3) o11_load_loop(L, B, c, P_, y[0]);
  }
  const double tmax = beats_end * bcl;
  const int first_win = B.nbeats_total - B.nwin;
  long long n = 0;
  bool active = valid && t < tmax && n < max_steps;
  while (__ballot(active) != 0ull) {
    ord11::Gates g;
    ord11::CaCurrents C;
    ord11::Currents cur;
    double f[ord11::NS];
    double dt = 0.0;
    // ---- rates and currents
    if (active) {
      if (wave == 0) {
        ord11::gate_rates_na<R>(y, P, g);
      } else if (wave == 1) {
        ord11::gate_rates_ca<R>(y, P, g);
      } else if (wave == 2) {
        ord11::NaKCurrents N;
        ord11::nak_currents<R>(y, P, N);
        s_nak[0][lane] = N.INa; s_nak[1][lane] = N.INaL; s_nak[2][lane] = N.Ito; s_nak[3][lane] = N.IKr;
        s_nak[4][lane] = N.IKs; s_nak[5][lane] = N.IK1; s_nak[6][lane] = N.IKb; s_nak[7][lane] = N.INab;
      } else {
        ord11::ca_currents<R>(y, P, C);
      }
    }
    __syncthreads();
    // ---- rates of V and the concentrations, step size, beat end
    if (active && wave == 3) {
      ord11::NaKCurrents N;
      N.INa = s_nak[0][lane]; N.INaL = s_nak[1][lane]; N.Ito = s_nak[2][lane]; N.IKr = s_nak[3][lane];
      N.IKs = s_nak[4][lane]; N.IK1 = s_nak[5][lane]; N.IKb = s_nak[6][lane]; N.INab = s_nak[7][lane];
      ord11::rates_from_currents<R>(t, y, f, P, cur, g, N, C);
      const double dts = ord11::time_step(t, B.time_point, B.max_dt, bcl, f[0]);
      if (floor((t + dts) / bcl) == floor(t / bcl)) {
        dt = dts;
      } else {
        dt = (floor(t / bcl) + 1) * bcl - t;
        o11_beat_end(L, B.bmd + c, P_, first_win, y[0]);
      }
      s_dt[lane] = dt;
    }
    __syncthreads();
    // ---- updates
    if (active) {
      if (wave != 3) dt = s_dt[lane];
      if (wave == 0) {
        ord11::rl_gates_na<R>(dt, y, g);
#pragma unroll
        for (int i = ord11::G0; i <= 24; ++i) s_gate[i - ord11::G0][lane] = y[i];
      } else if (wave == 1) {
        ord11::rl_gates_ca<R>(dt, y, g);
#pragma unroll
        for (int i = 25; i <= 38; ++i) s_gate[i - ord11::G0][lane] = y[i];
      } else if (wave == 3) {
        ord11::rl_core<R>(dt, y, f, g);
#pragma unroll
        for (int i = 0; i <= 9; ++i) s_core[i][lane] = y[i];
        o11_features(L, first_win, bcl, t, dt, y[0], f[0], cur);
      }
    }
    __syncthreads();
    // ---- every wave takes the new state it reads
    if (active) {
      if (wave == 0) {
        y[0] = s_core[0][lane];
      } else if (wave == 1) {
        y[0] = s_core[0][lane];
        y[2] = s_core[2][lane];
      } else {
#pragma unroll
        for (int i = 0; i <= 9; ++i) if (wave == 2) y[i] = s_core[i][lane];
#pragma unroll
        for (int i = ord11::G0; i <= 38; ++i)
          if (wave == 3 || i <= 24 || i >= 34) y[i] = s_gate[i - ord11::G0][lane];
      }
      ++n;
      t += dt;
      active = t < tmax && n < max_steps;
    }
  }
  if (!valid || wave != 3) return;
  o11_store(B, c, P_, y, L, t, tmax, n);
}

// The same loop with each cell's step split over the two wavefronts of a 128-thread workgroup
// (lane l of both holds cell blockIdx.x * 64 + l), on two SIMDs: wave 0 evaluates the gate
// kinetics (gate_rates, about half of the transcendentals) and their Rush-Larsen update;
// wave 1 the currents, the step size, the Jrel / Euler update and the features.  Per step
// they exchange the step size, then the new gates and V / cass, through LDS between two
// barriers.  Per cell the arithmetic is that of ord11_rl_kernel.  Both waves see the same t
// and step count for every lane, so they run the same number of passes of the loop.
template <typename R>
__global__ void __launch_bounds__(128) ord11_rl2_kernel(O11Buffers B, int beats_end, long long max_steps)
{
  constexpr int NGT = 38 - ord11::G0 + 1;       // gates 10..38 (incl. nca)
  __shared__ double s_gate[NGT][64];
  __shared__ double s_vc[2][64];
  __shared__ double s_dt[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  const bool valid = c < B.ncell && B.status[c] == ST_OK;
  const size_t P_ = (size_t)B.pitch;
  const PrmAcc P{B.prm + (valid ? c : 0), B.pitch};
  double y[ord11::NS];
  double st[O11_SD_N];
  double *sd = B.bmd + c;
  int *si = B.bmi + c;
  int pace = 0, elig = 0, steep = 0;
  double best_dvdt = -999.0, t = 0.0, bcl = 1.0;
  if (valid) {
#pragma unroll
    for (int i = 0; i < ord11::NS; ++i) y[i] = B.y[i * P_ + c];
#pragma unroll
    for (int k = 0; k < O11_SD_N; ++k) st[k] = sd[k * P_];
    bcl = P(ord11::P_BCL);
    t = st[O11_T];
    if (wave == 1) {
      pace = si[O11_I_PACE * P_]; elig = si[O11_I_ELIG * P_]; steep = si[O11_I_STEEP * P_];
      best_dvdt = sd[(O11_BEST + CEP_BM11_DVMDT_REPOL) * P_];
      if (si[O11_I_INIT * P_] == 0) {
        st[O11_TMP_DVDT] = -999.0; st[O11_TMP_PEAK] = -999.0; st[O11_TMP_VALLEY] = y[0];
        for (int k = 0; k < CEP_NBM; ++k) sd[(O11_BEST + k) * P_] = 0.0;
        sd[(O11_BEST + CEP_BM11_DVMDT_REPOL) * P_] = -999.0;
        sd[(O11_BEST + CEP_BM11_VM_PEAK) * P_] = -999.0;
        sd[(O11_BEST + CEP_BM11_VM_VALLEY) * P_] = y[0];
        best_dvdt = -999.0;
        si[O11_I_INIT * P_] = 1;
      }
    }
  }
  const double tmax = beats_end * bcl;
  const int first_win = B.nbeats_total - B.nwin;
  long long n = 0;
  bool active = valid && t < tmax && n < max_steps;
  while (__ballot(active) != 0ull) {
    ord11::Gates g;
    double f[ord11::NS];
    ord11::Currents cur;
    double dt = 0.0;
    if (wave == 0) {
      if (active) ord11::gate_rates<R>(y, P, g);
    } else if (active) {
      ord11::current_rates<R>(t, y, f, P, cur, g);
      const double dts = ord11::time_step(t, B.time_point, B.max_dt, bcl, f[0]);
      if (floor((t + dts) / bcl) == floor(t / bcl)) {
        dt = dts;
      } else {
        dt = (floor(t / bcl) + 1) * bcl - t;
        if (elig && pace >= first_win) {
          if (st[O11_TMP_DVDT] > best_dvdt) {
            best_dvdt = st[O11_TMP_DVDT];
            steep = pace;
            double *bs = sd + O11_BEST * P_;
            bs[CEP_BM11_QNET_AP * P_] = st[O11_QAP];
            bs[CEP_BM11_QNET4_AP * P_] = st[O11_Q4AP];
            bs[CEP_BM11_INAL_AUC_AP * P_] = st[O11_INALAP];
            bs[CEP_BM11_ICAL_AUC_AP * P_] = st[O11_ICALAP];
            bs[CEP_BM11_QNET_CL * P_] = st[O11_QCL];
            bs[CEP_BM11_QNET4_CL * P_] = st[O11_Q4CL];
            bs[CEP_BM11_INAL_AUC_CL * P_] = st[O11_INALCL];
            bs[CEP_BM11_ICAL_AUC_CL * P_] = st[O11_ICALCL];
            bs[CEP_BM11_VM_PEAK * P_] = st[O11_TMP_PEAK];
            bs[CEP_BM11_VM_VALLEY * P_] = st[O11_TMP_VALLEY];
            bs[CEP_BM11_DVMDT_REPOL * P_] = st[O11_TMP_DVDT];
            bs[CEP_BM11_T_PEAK * P_] = st[O11_TPEAK];
            bs[CEP_BM11_VM_REPOL30 * P_] = st[O11_R30];
            bs[CEP_BM11_VM_REPOL50 * P_] = st[O11_R50];
            bs[CEP_BM11_VM_REPOL90 * P_] = st[O11_R90];
          }
        }
#pragma unroll
        for (int k = O11_QAP; k <= O11_ICALCL; ++k) st[k] = 0.0;
        st[O11_TPEAK] = 0.0;
        st[O11_TMP_DVDT] = -999.0; st[O11_TMP_PEAK] = -999.0; st[O11_TMP_VALLEY] = y[0];
        pace++;
        elig = 0;
      }
      s_dt[lane] = dt;
    }
    __syncthreads();
    if (wave == 0) {
      if (active) {
        dt = s_dt[lane];
        ord11::rl_gates<R>(dt, y, g);
#pragma unroll
        for (int i = 0; i < NGT; ++i) s_gate[i][lane] = y[ord11::G0 + i];
      }
    } else if (active) {
      ord11::rl_core<R>(dt, y, f, g);
      s_vc[0][lane] = y[0];
      s_vc[1][lane] = y[2];
      if (pace >= first_win) {
        const double t_beat = bcl * pace + ord11::stim_start;
        if (t > t_beat + 2 && t < t_beat + 10 && fabs(cur.INa) < 1) {
          if (y[0] > st[O11_TMP_PEAK]) {
            st[O11_TMP_PEAK] = y[0];
            if (y[0] > 0) {
              const double amp = st[O11_TMP_PEAK] - st[O11_TMP_VALLEY];
              st[O11_R30] = st[O11_TMP_PEAK] - 0.3 * amp;
              st[O11_R50] = st[O11_TMP_PEAK] - 0.5 * amp;
              st[O11_R90] = st[O11_TMP_PEAK] - 0.9 * amp;
              elig = 1;
              st[O11_TPEAK] = t;
            } else {
              elig = 0;
            }
          }
        } else if (t > t_beat + 10 && elig) {
          if (f[0] > st[O11_TMP_DVDT] && y[0] <= st[O11_R30] && y[0] >= st[O11_R90]) st[O11_TMP_DVDT] = f[0];
        }
        const double inet = cur.INaL + cur.ICaL + cur.Ito + cur.IKr + cur.IKs + cur.IK1;
        const double inet4 = cur.INaL + cur.ICaL + cur.IKr + cur.INa;
        if (elig && y[0] > st[O11_R90]) {
          st[O11_QAP] += (inet * dt) / 1000.;
          st[O11_Q4AP] += (inet4 * dt) / 1000.;
          st[O11_INALAP] += cur.INaL * dt;
          st[O11_ICALAP] += cur.ICaL * dt;
        }
        st[O11_QCL] += (inet * dt) / 1000.;
        st[O11_Q4CL] += (inet4 * dt) / 1000.;
        st[O11_INALCL] += cur.INaL * dt;
        st[O11_ICALCL] += cur.ICaL * dt;
      }
    }
    __syncthreads();
    if (active) {
      if (wave == 0) {
        y[0] = s_vc[0][lane];
        y[2] = s_vc[1][lane];
      } else {
#pragma unroll
        for (int i = 0; i < NGT; ++i) y[ord11::G0 + i] = s_gate[i][lane];
      }
      ++n;
      t += dt;
      active = t < tmax && n < max_steps;
    }
  }
  if (!valid || wave == 0) return;
  st[O11_T] = t;
  bool fin = true;
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) { B.y[i * P_ + c] = y[i]; fin = fin && isfinite(y[i]); }
  if (!fin) B.status[c] = ST_NONFINITE;
  else if (t < tmax) B.status[c] = ST_MAX_STEPS;
#pragma unroll
  for (int k = 0; k < O11_SD_N; ++k) sd[k * P_] = st[k];
  si[O11_I_PACE * P_] = pace; si[O11_I_ELIG * P_] = elig; si[O11_I_STEEP * P_] = steep;
  B.stats[c] += n;              // steps (one right-hand side each)
}

// ---- ORd 2011 with the Adams PECE solver of the hackathon driver
// (CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:119-146): nout[c] further output points
// of dt each, every one an ode() call continuing the integration; optional Vm trace of the
// last trace_n outputs of the launch.
template <typename R>
__global__ void __launch_bounds__(64) ord11_adams_kernel(O11Buffers B, adams::Buffers A, int nbeats, int nout_fixed,
                                                        double dt, double relerr, double abserr,
                                                        double *vtrace, int tcap)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B.ncell) return;
  if (B.status[c] != ST_OK) return;
  const size_t P_ = (size_t)B.pitch;
  const PrmAcc P{B.prm + c, B.pitch};
  auto f = [&](double t, const double *y, double *yp) {
    ord11::Gates g;
    ord11::Currents cur;
    ord11::rhs<R>(t, y, yp, P, cur, g);
  };
  adams::Lane<ord11::NS, decltype(f)> L{A, (size_t)c, P_, f};
  L.load();
  double y[ord11::NS];
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) y[i] = B.y[i * P_ + c];
  double t = A.sc[adams::S_T * P_ + c];
  int iflag = A.si[adams::I_IFLAG * P_ + c];
  // outputs of this launch: nbeats beats of the cell's own BCL, or a fixed count
  const int nout = nout_fixed > 0 ? nout_fixed : nbeats * (int)llround(P(ord11::P_BCL) / dt);
  const int trace_from = nout - tcap;
  int st = ST_OK;
  for (int o = 0; o < nout; ++o) {
    const double tout = t + dt;
    iflag = L.de(y, t, tout, relerr, abserr, iflag == 0 ? 1 : iflag);
    if (iflag != 2) {
      st = (iflag == 4 || iflag == 5 || iflag == -4 || iflag == -5) ? ST_MAX_STEPS : ST_ERR_FAIL;
      break;
    }
    if (vtrace && o >= trace_from) vtrace[(size_t)(o - trace_from) * P_ + c] = y[0];
  }
  bool fin = true;
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) { B.y[i * P_ + c] = y[i]; fin = fin && isfinite(y[i]); }
  if (st == ST_OK && !fin) st = ST_NONFINITE;
  if (st != ST_OK) B.status[c] = st;
  A.sc[adams::S_T * P_ + c] = t;
  A.si[adams::I_IFLAG * P_ + c] = iflag;
  L.store();
}

// computeRates of ORd 2011 for a batch (parity path): AoS states in, AoS rates and the
// ALGEBRAIC current slots out
constexpr int ORD11_NALG = 198;
template <typename R>
__global__ void __launch_bounds__(64) ord11_rates_kernel(int n, int pitch, double t, const double *__restrict__ yin,
                                                        double *__restrict__ fout, double *__restrict__ alg,
                                                        const double *__restrict__ prm)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double y[ord11::NS], f[ord11::NS];
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) y[i] = yin[(size_t)c * ord11::NS + i];
  const PrmAcc P{prm + c, pitch};
  ord11::Gates g;
  ord11::Currents cur;
  ord11::rhs<R>(t, y, f, P, cur, g);
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) fout[(size_t)c * ord11::NS + i] = f[i];
  if (alg) {
    double *a = alg + (size_t)c * ORD11_NALG;
    a[12] = cur.Istim; a[62] = cur.INa; a[64] = cur.INaL; a[70] = cur.Ito; a[81] = cur.ICaL;
    a[96] = cur.IKr; a[98] = cur.IKs; a[100] = cur.IK1;
  }
}

// ============================================================================ host side

namespace host {

// ---- ORd-CiPA CellML constants and initial states (___initConsts, :545-804)
static const double kOrdInitStates[49] = {
  -88.00190465, 0.0125840447, 8.49e-05, 7.268004498, 7.268089977, 144.6555918, 144.6555651,
  1.619574538, 1.571234014, 8.6e-05, 0.007344121102, 0.6981071913, 0.6980895801, 0.6979908432,
  0.4549485525, 0.6979245865, 0.0001882617273, 0.5008548855, 0.2693065357, 0.001001097687,
  0.9995541745, 0.5865061736, 0.0005100862934, 0.9995541823, 0.6393399482, 2.34e-9,
  0.9999999909, 0.9102412777, 0.9999999909, 0.9998046777, 0.9999738312, 0.9999999909,
  0.9999999909, 0.002749414044, 0.999637, 6.83208e-05, 1.80145e-08, 8.26619e-05, 0.00015551,
  5.67623e-05, 0, 0, 0, 0, 0.2707758025, 0.0001928503426, 0.9967597594, 2.5e-7, 3.12e-7 };

// Markov IKr (A, B, q) per transition, order k1 k2 k3 k4 k11 k21 k31 k41 k51 k52 k53 k61 k62 k63
static const double kMarkovABq[14][3] = {
  { 0.0264, 4.631E-05, 4.843 }, { 4.986E-06, -0.004226, 4.23 }, { 0.001214, 0.008516, 4.962 },
  { 1.854E-05, -0.04641, 3.769 }, { 0.0007868, 1.535E-08, 4.942 }, { 5.455E-06, -0.1688, 4.156 },
  { 0.005509, 7.771E-09, 4.22 }, { 0.001416, -0.02877, 1.459 }, { 0.4492, 0.008595, 5 },
  { 0.3181, 3.613E-08, 4.663 }, { 0.149, 0.004668, 2.412 }, { 0.01241, 0.1725, 5.568 },
  { 0.3226, -0.0006575, 5 }, { 0.008978, -0.02215, 5.682 } };

static double hill_factor(double D, double ic50, double h)
{
  return (ic50 > 10E-14 && h > 10E-14) ? pow(1. + pow(D / ic50, h), -1) : 1.;
}

struct OrdCellIn {
  double celltype, conc, bcl;
  double herg[19];
  double gs[7];
  double kt;            // drug trapping rate Kt, CONSTANTS[100] (0 in the reference's ___initConsts)
};

// per-cell parameters (celltype-derived constants :749-803, Hill block :826-842,
// population scales applied after initConsts)
static void ord_params(const OrdCellIn &in, double stim_scale, double *p)
{
  using namespace ord;
  const bool epi = in.celltype == 1.0, mid = in.celltype == 2.0;
  const double D = in.conc;
  const double *hg = in.herg;
  const double pca_base = epi ? 0.0001007 * 1.2 : mid ? 0.0001007 * 2.5 : 0.0001007;
  const double pcap = 1.1 * pca_base;
  p[OP_CELLTYPE] = in.celltype;
  p[OP_GNA] = 75.0 * hill_factor(D, hg[6], hg[7]) * in.gs[0];
  p[OP_GNAL] = (epi ? 0.019957499999999975 * 0.6 : 0.019957499999999975) * hill_factor(D, hg[8], hg[9]) * in.gs[1];
  p[OP_PCA] = pca_base * hill_factor(D, hg[0], hg[1]) * in.gs[2];
  p[OP_PCAP] = pcap;
  p[OP_PCANA] = 0.00125 * pca_base;
  p[OP_PCAK] = 0.0003574 * pca_base;
  p[OP_PCANAP] = 0.00125 * pcap;
  p[OP_PCAKP] = 0.0003574 * pcap;
  p[OP_GKR] = (epi ? 0.04658545454545456 * 1.3 : mid ? 0.04658545454545456 * 0.8 : 0.04658545454545456) * in.gs[3];
  p[OP_GKS] = (epi ? 0.006358000000000001 * 1.4 : 0.006358000000000001) * hill_factor(D, hg[4], hg[5]) * in.gs[4];
  p[OP_GK1] = (epi ? 0.3239783999999998 * 1.2 : mid ? 0.3239783999999998 * 1.3 : 0.3239783999999998) * hill_factor(D, hg[2], hg[3]) * in.gs[5];
  p[OP_GTO] = ((epi || mid) ? 0.02 * 4.0 : 0.02) * hill_factor(D, hg[10], hg[11]) * in.gs[6];
  p[OP_GKB] = epi ? 0.003 * 0.6 : 0.003;
  p[OP_CMDNMAX] = epi ? 0.05 * 1.3 : 0.05;
  p[OP_UPSCALE] = epi ? 1.3 : 1.0;
  p[OP_GNCX] = epi ? 0.0008 * 1.1 : mid ? 0.0008 * 1.4 : 0.0008;
  p[OP_PNAK] = epi ? 30 * 0.9 : mid ? 30 * 0.7 : 30;
  const double Kmax = hg[14], Ku = hg[15], n = hg[16], halfmax = hg[17];
  p[OP_KU] = Ku;
  p[OP_BIND] = (Kmax * Ku * exp(n * log(D))) / (exp(n * log(D)) + halfmax);
  p[OP_KT] = in.kt;
  p[OP_VHALF] = hg[18];
  p[OP_STIM_START] = 3;
  p[OP_STIM_END] = 100000000000000000.0;
  p[OP_STIM_AMP] = -80 * stim_scale;
  p[OP_BCL] = in.bcl;
  p[OP_STIM_DUR] = 0.5;
  p[OP_CONC] = D;
  p[OP_JREL_SCALE] = mid ? 1.7 : 1.0;
  p[OP_EPI] = epi ? 1.0 : 0.0;
}

// Full CellML CONSTANTS[206] of a cell, for hosts that read the reference layout
static void ord_constants(const OrdCellIn &in, double stim_scale, double *C)
{
  memset(C, 0, sizeof(double) * 206);
  double p[ord::NP];
  ord_params(in, stim_scale, p);
  const double tab[][2] = {
    { 1, 140 }, { 2, 1.8 }, { 3, 5.4 }, { 4, 8314 }, { 5, 310 }, { 6, 96485 }, { 7, 1 }, { 8, 2 }, { 9, 1 },
    { 10, 0.01 }, { 11, 0.0011 }, { 17, 0.15 }, { 18, 0.05 }, { 19, 0.00068 }, { 20, 0.05 }, { 21, 0.0015 },
    { 22, 0.05 }, { 23, 0.00238 }, { 24, 0.07 }, { 25, 0.0005 }, { 26, 0.047 }, { 27, 0.00087 }, { 28, 1.124 },
    { 29, 0.0087 }, { 30, 10 }, { 31, 0.8 }, { 32, 1 }, { 33, 0.01833 }, { 34, 39.57 }, { 35, 9.871 },
    { 36, 11.64 }, { 37, 34.77 }, { 38, 6.765 }, { 39, 8.552 }, { 40, 77.42 }, { 41, 5.955 }, { 42, 82.9 },
    { 43, 6.086 }, { 44, 0.99 }, { 47, 200 }, { 48, 0.019957499999999975 }, { 49, 0.02 }, { 50, 0.002 },
    { 51, 1000 }, { 52, 0.0001007 }, { 53, 0.04658545454545456 }, { 102, 37 }, { 103, 0.006358000000000001 },
    { 104, 817.3 }, { 105, 0.3239783999999998 }, { 106, 15 }, { 107, 5 }, { 108, 88.12 }, { 109, 12.5 },
    { 110, 6e4 }, { 111, 6e4 }, { 112, 5e3 }, { 113, 1.5e6 }, { 114, 5e3 }, { 115, 0.5224 }, { 116, 0.167 },
    { 117, 150e-6 }, { 118, 0.0008 }, { 119, 949.5 }, { 120, 182.4 }, { 121, 687.2 }, { 122, 39.4 },
    { 123, 1899 }, { 124, 79300 }, { 125, 639 }, { 126, 40 }, { 127, 9.073 }, { 128, 27.78 }, { 129, -0.155 },
    { 130, 0.5 }, { 131, 0.3582 }, { 132, 0.05 }, { 133, 9.8 }, { 134, 1.698e-7 }, { 135, 1e-7 }, { 136, 4.2 },
    { 137, 1.698e-7 }, { 138, 224 }, { 139, 292 }, { 140, 30 }, { 141, 0.003 }, { 142, 3.75e-10 },
    { 143, 2.5e-8 }, { 144, 0.0005 }, { 145, 0.0005 }, { 146, 4.75 }, { 147, 1.0 }, { 148, 1.0 } };
  for (const auto &e : tab) C[(int)e[0]] = e[1];
  for (int k = 0; k < 14; ++k) for (int j = 0; j < 3; ++j) C[54 + 3 * k + j] = 0;
  {
    // CellML order of the Markov parameter triplets at CONSTANTS[54..95]
    const int order[14] = { 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13 };
    for (int k = 0; k < 14; ++k) for (int j = 0; j < 3; ++j) C[54 + 3 * k + j] = kMarkovABq[order[k]][j];
  }
  using namespace ord;
  C[0] = in.celltype;
  C[12] = p[OP_STIM_START]; C[13] = p[OP_STIM_END]; C[14] = p[OP_STIM_AMP]; C[15] = p[OP_BCL]; C[16] = p[OP_STIM_DUR];
  C[45] = p[OP_GNA];
  C[96] = in.herg[14]; C[97] = in.herg[15]; C[98] = in.herg[16]; C[99] = in.herg[17]; C[100] = in.kt; C[101] = in.herg[18];
  C[149] = frt; C[150] = p[OP_CMDNMAX]; C[151] = Ahs; C[152] = thLp; C[153] = p[OP_GNAL]; C[154] = p[OP_GTO];
  C[155] = 0.6; C[156] = p[OP_PCA]; C[157] = 75; C[158] = 0; C[159] = p[OP_GKR]; C[160] = p[OP_GKS];
  C[161] = p[OP_GK1]; C[162] = vcell; C[163] = p[OP_GKB]; C[164] = 0; C[165] = 0; C[166] = a_rel; C[167] = btp;
  C[168] = p[OP_UPSCALE]; C[169] = ffrt; C[170] = 1 - 0.6; C[171] = p[OP_PCAP]; C[172] = p[OP_PCANA];
  C[173] = p[OP_PCAK]; C[174] = 2 * frt; C[175] = frt; C[176] = frt; C[177] = Ageo; C[178] = frt;
  C[179] = 2 * frt; C[180] = a_relp; C[181] = p[OP_PCANAP]; C[182] = p[OP_PCAKP]; C[183] = Acap;
  C[184] = vmyo; C[185] = vnsr; C[186] = vjsr; C[187] = vss; C[188] = h10; C[189] = h11; C[190] = h12;
  C[191] = k1ncx; C[192] = kcaoff; C[193] = kcaoff; C[194] = p[OP_GNCX]; C[195] = h10; C[196] = h11;
  C[197] = h12; C[198] = k1ncx; C[199] = kcaoff; C[200] = kcaoff; C[201] = b1nak; C[202] = a2nak;
  C[203] = a4nak; C[204] = p[OP_PNAK]; C[205] = 0;
}

// ---- TT06 constants / initial states (initConsts :196-270 and the epi/M siblings)
static const double kTtInit[3][19] = {
  { -86.709, 138.4, 10.355, 0.00013, 0.00448, 0.476, 0.0087, 0.00155, 0.7573, 0.7225, 0.00036,
    3.164e-5, 0.8009, 0.9778, 0.9953, 0.3212, 2.235e-8, 3.715, 0.9068 },
  { -85.23, 136.89, 8.604, 0.000126, 0.00621, 0.4712, 0.0095, 0.00172, 0.7444, 0.7045, 0.00036,
    3.373e-5, 0.7888, 0.9755, 0.9953, 0.999998, 2.42e-8, 3.64, 0.9073 },
  { -85.423, 138.52, 10.132, 0.000153, 0.0165, 0.473, 0.0174, 0.00165, 0.749, 0.6788, 0.00042,
    3.288e-5, 0.7026, 0.9526, 0.9942, 0.999998, 2.347e-8, 4.272, 0.8978 } };

struct TtCellIn {
  double celltype, conc, bcl;
  double ic50[3], hill[3];
  double gs[10];
};

static void tt_params(const TtCellIn &in, double stim_scale, double *p)
{
  using namespace tt;
  const int ct = (int)in.celltype;
  const double base[10] = { 14.838, 5.405, 0.153, ct == 2 ? 0.098 : 0.392, ct == 0 ? 0.073 : 0.294,
                            0.0000398, 0.00029, 0.000592, 0.1238, 0.0146 };
  const int slot[10] = { TP_GNA, TP_GK1, TP_GKR, TP_GKS, TP_GTO, TP_GCAL, TP_GBNA, TP_GBCA, TP_GPCA, TP_GPK };
  p[TP_CELLTYPE] = in.celltype;
  for (int k = 0; k < 10; ++k) p[slot[k]] = base[k] * in.gs[k];
  // Hill block on IKr, ICaL, INa with the CiPA initConsts rule
  p[TP_GKR] *= hill_factor(in.conc, in.ic50[0], in.hill[0]);
  p[TP_GCAL] *= hill_factor(in.conc, in.ic50[1], in.hill[1]);
  p[TP_GNA] *= hill_factor(in.conc, in.ic50[2], in.hill[2]);
  p[TP_PNAK] = 2.724;
  p[TP_STIM_START] = 10;
  p[TP_BCL] = in.bcl;
  p[TP_STIM_DUR] = 1;
  p[TP_STIM_AMP] = 52 * stim_scale;
}

static void tt_constants(const TtCellIn &in, double stim_scale, double *C)
{
  static const double c0[53] = {
    8314.472, 310, 96485.3415, 0.185, 0.016404, 10, 1000, 1, 52, 0.03, 5.4, 140, 2, 5.405, 0.153,
    0.392, 14.838, 0.00029, 0.0000398, 0.000592, 0.073, 2.724, 1, 40, 1000, 0.1, 2.5, 0.35, 1.38,
    87.5, 0.1238, 0.0005, 0.0146, 0.15, 0.045, 0.06, 0.005, 1.5, 2.5, 1, 0.102, 0.0038, 0.00025,
    0.00036, 0.006375, 0.2, 0.001, 10, 0.3, 0.4, 0.00025, 0.001094, 0.00005468 };
  memcpy(C, c0, sizeof(c0));
  double p[tt::NP];
  tt_params(in, stim_scale, p);
  using namespace tt;
  C[16] = p[TP_GNA]; C[13] = p[TP_GK1]; C[14] = p[TP_GKR]; C[15] = p[TP_GKS]; C[20] = p[TP_GTO];
  C[18] = p[TP_GCAL]; C[17] = p[TP_GBNA]; C[19] = p[TP_GBCA]; C[30] = p[TP_GPCA]; C[32] = p[TP_GPK];
  C[6] = p[TP_BCL]; C[8] = p[TP_STIM_AMP];
}

// ORd 2011 (DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp): O11CellIn, kO11Init and o11_params
// are in ord11_host.hpp

// CONSTANTS[144] in the reference layout (enum E_CONSTANTS_T)
static void o11_constants(const O11CellIn &in, bool dutta, double stim_scale, double *C)
{
  memset(C, 0, sizeof(double) * ord11::NC);
  static const double prim[93] = {
    0, 140, 1.8, 5.4, 8314, 310, 96485, 1, 2, 1, 0.01, 0.0011, -80, 0.5, 0.15, 0.05, 0.00068, 0.05,
    0.0015, 0.05, 0.00238, 0.07, 0.0005, 0.047, 0.00087, 1.124, 0.0087, 10, 0.8, 1, 0.01833, 39.57,
    9.871, 11.64, 34.77, 6.765, 8.552, 77.42, 5.955, 82.9, 6.086, 0.99, 75, 200, 0.0075, 0.02, 0.002,
    1000, 0.0001, 0.046, 0.0034, 0.1908, 15, 5, 88.12, 12.5, 6e4, 6e4, 5e3, 1.5e6, 5e3, 0.5224, 0.167,
    150e-6, 0.0008, 949.5, 182.4, 687.2, 39.4, 1899, 79300, 639, 40, 9.073, 27.78, -0.155, 0.5, 0.3582,
    0.05, 9.8, 1.698e-7, 1e-7, 4.2, 1.698e-7, 224, 292, 30, 0.003, 3.75e-10, 2.5e-8, 0.0005, 0.0005, 4.75 };
  memcpy(C, prim, sizeof(prim));
  using namespace ord;
  double p[ord11::NP];
  o11_params(in, dutta, stim_scale, p);
  C[12] = p[ord11::P_STIM_AMP];
  C[42] = p[ord11::P_GNA];                     // GNa with the drug block
  C[93] = 0.05; C[94] = 1.0 - 0.99; C[95] = 3.0 * 200; C[96] = p[ord11::P_GNAL]; C[97] = p[ord11::P_GTO];
  C[98] = p[ord11::P_AFF]; C[99] = p[ord11::P_PCA]; C[100] = p[ord11::P_TJCA]; C[101] = p[ord11::P_GKR];
  C[102] = p[ord11::P_GKS]; C[103] = p[ord11::P_GK1]; C[104] = vcell; C[105] = 0.003; C[106] = a_rel;
  C[107] = btp; C[108] = 1.0; C[109] = 1.0 - 0.6; C[110] = ord11::PCap; C[111] = ord11::PCaNa;
  C[112] = ord11::PCaK; C[113] = Ageo; C[114] = a_relp; C[115] = ord11::PCaNap; C[116] = ord11::PCaKp;
  C[117] = Acap; C[118] = vmyo; C[119] = vnsr; C[120] = vjsr; C[121] = vss; C[122] = h10; C[123] = h11;
  C[124] = h12; C[125] = k1ncx; C[126] = kcaoff; C[127] = kcaoff; C[128] = 0.0008; C[129] = h10;
  C[130] = h11; C[131] = h12; C[132] = k1ncx; C[133] = kcaoff; C[134] = kcaoff; C[135] = b1nak;
  C[136] = a2nak; C[137] = a4nak; C[138] = 30; C[139] = ord11::stim_start; C[140] = in.bcl;
  if (dutta) { C[81] = p[ord11::P_H]; C[64] = 0.0008 * 2.661; }
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
};

}  // namespace host
}  // namespace cep

using namespace cep;

struct cep_batch {
  int model, n, prec, pitch;
  int ns, nc, nalg, np;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ktime_total = 0.0;
  int klaunches = 0;
  bool pending = false;
  // run config
  double dt = 0.1, rtol = 1e-6, atol = 1e-8, stim_scale = 1.0;
  int maxord = 5, nbeats_total = 1, nwin = 1;
  bool vtrace_on = false;
  int tcap = 0, cap = 0;
  // time series of the selected beat (ORd-CiPA): [t, vm, ca][2][ts_cap], ires[6][2][ts_cap]
  bool ts_on = false;
  int ts_cap = 0, ts_nwin = 0;
  double *d_ts = nullptr, *d_tsv = nullptr;
  int *d_tsp = nullptr;
  // host copies of inputs
  std::vector<host::OrdCellIn> ord_in;
  std::vector<host::TtCellIn> tt_in;
  std::vector<host::O11CellIn> o11_in;
  bool dutta = false;
  double time_point = 25.0;
  int beats_done = 0;
  int integrator = CEP_INTEGRATOR_RUSH_LARSEN;
  int o11_waves = 0;             // CEP_OPT_ORD11_WAVES: 0 automatic, 1, 2 or 4 wavefronts per 64 cells
  // launch shapes (CEP_SHAPE_*, 0 = automatic), chosen by hand, by cep_autotune or from its cache
  int shape_wpe = 0, shape_block = 0, shape_order = 0;
  int *d_gorder = nullptr;       // ORd-CiPA: cell group of each workgroup, longest beat first
  std::vector<int> h_gorder;
  double *d_aphi = nullptr, *d_avec = nullptr, *d_acoef = nullptr, *d_asc = nullptr;
  int *d_asi = nullptr;
  // biomarker rules of the drugs_herg / benchpatch / population_tn2006 drivers
  bool dr_on = false;
  double v_apd90 = -40.0, v_apd50 = -40.0, t_first = 0.0;   // t_first 0: the first output at dt
  double *d_dr = nullptr, *d_qin = nullptr;
  int *d_qctl = nullptr;
  // device memory
  double *d_y = nullptr, *d_prm = nullptr, *d_bmd = nullptr, *d_ct = nullptr, *d_cv = nullptr;
  double *d_zn = nullptr, *d_ewt = nullptr, *d_acor = nullptr, *d_ft = nullptr, *d_jd = nullptr;
  double *d_fac = nullptr, *d_sd = nullptr, *d_tcur = nullptr, *d_vtrace = nullptr;
  int *d_si = nullptr, *d_bmi = nullptr, *d_status = nullptr, *d_iout = nullptr;
  long long *d_stats = nullptr;
  // cells restarted one by one (cep_restart_cells): each cell's first window beat (-1: the
  // batch's) and the staging rows of a restart
  int *d_win = nullptr;
  double *d_stage = nullptr;
  int *d_stage_idx = nullptr;
  std::vector<double> h_stage;
  std::vector<int> h_stage_idx;
  std::vector<std::pair<void *, size_t>> allocs;   // device blocks and their sizes
};

// Device memory of batches is recycled, not freed: hipFree waits for the work of every stream
// on the device, so in a serving loop the end of one job would hold the host until every
// other job's running beat kernel had finished.  A destroyed batch's blocks (its stream
// drained first) go to a free list per device and are handed to later batches; the smallest
// block of at least the request and at most twice it is taken.  cep_release_cached_memory()
// returns the list to the device.
namespace blockcache {
struct Block {
  void *p;
  size_t bytes;
  int dev;
};
static std::mutex mu;
static std::vector<Block> cached;

static size_t round_bytes(size_t n) { return (n + 65535) & ~(size_t)65535; }

static hipError_t get(void **p, size_t bytes, size_t *got)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  bytes = round_bytes(bytes);
  {
    std::lock_guard<std::mutex> lk(mu);
    size_t best = cached.size();
    for (size_t i = 0; i < cached.size(); ++i) {
      const Block &c = cached[i];
      if (c.dev == dev && c.bytes >= bytes && c.bytes <= 2 * bytes &&
          (best == cached.size() || c.bytes < cached[best].bytes))
        best = i;
    }
    if (best < cached.size()) {
      *p = cached[best].p;
      *got = cached[best].bytes;
      cached.erase(cached.begin() + best);
      return hipSuccess;
    }
  }
  *got = bytes;
  return hipMalloc(p, bytes);
}

static void put(void *p, size_t bytes)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  cached.push_back(Block{ p, bytes, dev });
}
}  // namespace blockcache

static int fail(cep_batch *b, int code, const std::string &msg)
{
  if (b) b->err = msg;
  return code;
}

#define HIPCHK(b, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return fail(b, CEP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int dalloc(cep_batch *b, T **p, size_t count)
{
  void *q = nullptr;
  size_t got = 0;
  hipError_t e = blockcache::get(&q, count * sizeof(T) + 64, &got);
  if (e != hipSuccess) return fail(b, CEP_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back({ q, got });
  // on the batch stream: a memset on the null stream is not ordered with the batch's
  // (non-blocking) stream and could land after the first uploads into the buffer
  e = hipMemsetAsync(q, 0, count * sizeof(T) + 64, b->stream);
  if (e != hipSuccess) return fail(b, CEP_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  *p = (T *)q;
  return 0;
}

static int ensure_markov_table()
{
  static bool done = false;
  if (done) return 0;
  ord::MarkovTable t;
  for (int k = 0; k < 14; ++k) {
    const double A = host::kMarkovABq[k][0], Bc = host::kMarkovABq[k][1], q = host::kMarkovABq[k][2];
    t.AQ[k] = A * exp(((37.0 - 20.0) * log(q)) / 10.0);
    t.B[k] = Bc;
  }
  if (hipMemcpyToSymbol(HIP_SYMBOL(ord::g_markov), &t, sizeof(t)) != hipSuccess) return CEP_E_HIP;
  // batches run on non-blocking streams: make the table visible before any of them launches
  if (hipDeviceSynchronize() != hipSuccess) return CEP_E_HIP;
  done = true;
  return 0;
}

static int alloc_trace_buffers(cep_batch *b);

extern "C" {

// integrator = false: parameters and states only (no BDF workspace), for the tissue nodes
static cep_batch *batch_create(int model, int ncells, int precision, bool integrator);

cep_batch *cep_batch_create(int model, int ncells, int precision)
{
  return batch_create(model, ncells, precision, true);
}

}  // extern "C"

static cep_batch *batch_create(int model, int ncells, int precision, bool integrator)
{
  if ((model != CEP_MODEL_ORD_CIPA && model != CEP_MODEL_TT06 && model != CEP_MODEL_ORD2011) || ncells <= 0 ||
      (precision != CEP_PREC_FP64 && precision != CEP_PREC_MIXED))
    return nullptr;
  cep_batch *b = new cep_batch();
  b->model = model;
  b->n = ncells;
  b->prec = precision;
  b->pitch = (ncells + 63) / 64 * 64;
  if (model == CEP_MODEL_ORD_CIPA) {
    b->ns = ord::NS; b->nc = ord::NC; b->nalg = ORD_NALG; b->np = ord::NP;
    b->ord_in.resize(ncells);
    for (auto &c : b->ord_in) { memset(&c, 0, sizeof(c)); c.bcl = 1000; for (double &g : c.gs) g = 1.0; c.herg[16] = 1.0; c.herg[17] = 1.0; }
  } else if (model == CEP_MODEL_ORD2011) {
    b->ns = ord11::NS; b->nc = ord11::NC; b->nalg = ORD11_NALG; b->np = ord11::NP;
    b->dt = 1.0;        // max_time_step
    b->o11_in.resize(ncells);
    for (auto &c : b->o11_in) { memset(&c, 0, sizeof(c)); c.bcl = 1000; }
  } else {
    b->ns = tt::NS; b->nc = tt::NC; b->nalg = TT_NALG; b->np = tt::NP;
    b->dt = 0.02;
    b->tt_in.resize(ncells);
    for (auto &c : b->tt_in) { memset(&c, 0, sizeof(c)); c.bcl = 1000; for (double &g : c.gs) g = 1.0; }
  }
  const size_t P = b->pitch;
  int rc = 0;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    delete b;
    return nullptr;
  }
  rc |= dalloc(b, &b->d_y, P * b->ns);
  rc |= dalloc(b, &b->d_prm, P * b->np);
  rc |= dalloc(b, &b->d_bmd, P * BM_TOTAL_D);
  rc |= dalloc(b, &b->d_bmi, P * BM_I_N);
  rc |= dalloc(b, &b->d_status, P);
  rc |= dalloc(b, &b->d_win, P);
  if (model == CEP_MODEL_ORD_CIPA && integrator) {
    rc |= dalloc(b, &b->d_zn, P * 6 * ordbdf::N);
    rc |= dalloc(b, &b->d_ewt, P * ordbdf::N);
    rc |= dalloc(b, &b->d_acor, P * ordbdf::N);
    rc |= dalloc(b, &b->d_ft, P * ordbdf::N);
    rc |= dalloc(b, &b->d_jd, P * ordbdf::NJC * ordbdf::N);
    rc |= dalloc(b, &b->d_fac, P * ordbdf::NFAC);
    rc |= dalloc(b, &b->d_sd, P * ordbdf::SD_N);
    rc |= dalloc(b, &b->d_si, P * ordbdf::SI_N);
    rc |= dalloc(b, &b->d_stats, P * 8);
  } else if (model == CEP_MODEL_TT06) {
    rc |= dalloc(b, &b->d_tcur, P);
    rc |= dalloc(b, &b->d_iout, P);
  } else if (model == CEP_MODEL_ORD2011) {
    rc |= dalloc(b, &b->d_stats, 2 * P);    // steps, right-hand sides
  }
  if (rc || ensure_markov_table()) {
    cep_batch_destroy(b);
    return nullptr;
  }
  return b;
}

extern "C" {

int cep_release_cached_memory(void)
{
  std::lock_guard<std::mutex> lk(blockcache::mu);
  int rc = 0;
  for (const auto &c : blockcache::cached) rc |= hipFree(c.p) != hipSuccess;
  blockcache::cached.clear();
  return rc ? CEP_E_HIP : 0;
}

void cep_batch_destroy(cep_batch *b)
{
  if (!b) return;
  // the stream drained, no kernel of this batch touches its blocks any more
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto &a : b->allocs) blockcache::put(a.first, a.second);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *cep_last_error(const cep_batch *b) { return b ? b->err.c_str() : "null batch"; }

int cep_batch_info(const cep_batch *b, int *model, int *ncells, int *nstates, int *nconsts, int *nalg)
{
  if (!b) return CEP_E_ARG;
  if (model) *model = b->model;
  if (ncells) *ncells = b->n;
  if (nstates) *nstates = b->ns;
  if (nconsts) *nconsts = b->nc;
  if (nalg) *nalg = b->nalg;
  return 0;
}

int cep_set_cells(cep_batch *b, int first, int count, const double *celltype, const double *conc,
                  const double *herg, int herg_stride, const double *gscale, int gscale_stride,
                  const double *bcl)
{
  if (!b || first < 0 || count < 0 || first + count > b->n) return fail(b, CEP_E_ARG, "cell range");
  for (int k = 0; k < count; ++k) {
    const int i = first + k;
    if (b->model == CEP_MODEL_ORD_CIPA) {
      host::OrdCellIn &c = b->ord_in[i];
      if (celltype) c.celltype = celltype[k];
      if (conc) c.conc = conc[k];
      if (bcl) c.bcl = bcl[k];
      if (herg) for (int j = 0; j < 19; ++j) c.herg[j] = herg[(size_t)k * herg_stride + j];
      if (gscale) for (int j = 0; j < 7; ++j) c.gs[j] = gscale[(size_t)k * gscale_stride + j];
      if (!(c.bcl > 0) || c.celltype < 0 || c.celltype > 2) return fail(b, CEP_E_ARG, "bad celltype/bcl");
    } else if (b->model == CEP_MODEL_ORD2011) {
      host::O11CellIn &c = b->o11_in[i];
      if (gscale) return fail(b, CEP_E_ARG, "ORd 2011 takes no population scales");
      if (herg && herg_stride < 14) return fail(b, CEP_E_ARG, "ORd 2011 IC50 row needs 14 values");
      if (celltype) c.celltype = celltype[k];
      if (conc) c.conc = conc[k];
      if (bcl) c.bcl = bcl[k];
      if (herg) for (int j = 0; j < 14; ++j) c.ic50[j] = herg[(size_t)k * herg_stride + j];
      if (!(c.bcl > 0) || c.celltype < 0 || c.celltype > 2) return fail(b, CEP_E_ARG, "bad celltype/bcl");
    } else {
      host::TtCellIn &c = b->tt_in[i];
      if (celltype) c.celltype = celltype[k];
      if (conc) c.conc = conc[k];
      if (bcl) c.bcl = bcl[k];
      if (herg) for (int j = 0; j < 3; ++j) { c.ic50[j] = herg[(size_t)k * herg_stride + j]; c.hill[j] = herg[(size_t)k * herg_stride + 3 + j]; }
      if (gscale) for (int j = 0; j < 10; ++j) c.gs[j] = gscale[(size_t)k * gscale_stride + j];
      if (!(c.bcl > 0) || c.celltype < 0 || c.celltype > 2) return fail(b, CEP_E_ARG, "bad celltype/bcl");
    }
  }
  return 0;
}

int cep_set_run_config(cep_batch *b, double dt, double rtol, double atol, int maxord, int nbeats_total,
                       int nwin, double stim_amp_scale)
{
  if (!b || !(dt > 0) || !(rtol > 0) || !(atol > 0) || maxord < 1 || maxord > 5 || nbeats_total < 1 || nwin < 1)
    return fail(b, CEP_E_ARG, "run config");
  b->dt = dt; b->rtol = rtol; b->atol = atol; b->maxord = maxord; b->nbeats_total = nbeats_total;
  b->nwin = nwin; b->stim_scale = stim_amp_scale;
  return 0;
}

int cep_set_model_option(cep_batch *b, int option, double value)
{
  if (!b) return CEP_E_ARG;
  if (option == CEP_OPT_DRIVER_RULES || option == CEP_OPT_V_APD90) {
    if (b->model != CEP_MODEL_ORD_CIPA && b->model != CEP_MODEL_TT06)
      return fail(b, CEP_E_ARG, "driver rules are evaluated for ORd-CiPA and TT06");
    if (option == CEP_OPT_DRIVER_RULES) b->dr_on = value != 0.0;
    else b->v_apd90 = value;
    return 0;
  }
  if (option == CEP_OPT_V_APD50 || option == CEP_OPT_FIRST_OUTPUT) {
    if (b->model != CEP_MODEL_ORD_CIPA) return fail(b, CEP_E_ARG, "an ORd-CiPA option");
    if (option == CEP_OPT_V_APD50) b->v_apd50 = value;
    else if (value >= 0.0) b->t_first = value;
    else return fail(b, CEP_E_ARG, "first output time");
    return 0;
  }
  if (option == CEP_OPT_KT) {
    if (b->model != CEP_MODEL_ORD_CIPA) return fail(b, CEP_E_ARG, "Kt is an ORd-CiPA constant");
    if (!(value >= 0.0)) return fail(b, CEP_E_ARG, "Kt must be >= 0");
    for (auto &c : b->ord_in) c.kt = value;
    return 0;
  }
  if (option == CEP_OPT_TIME_SERIES) {
    if (b->model != CEP_MODEL_ORD_CIPA) return fail(b, CEP_E_ARG, "time series are recorded for ORd-CiPA");
    b->ts_on = value != 0.0;
    return 0;
  }
  if (b->model != CEP_MODEL_ORD2011) return fail(b, CEP_E_ARG, "no options for this model");
  if (option == CEP_OPT_DUTTA) b->dutta = value != 0.0;
  else if (option == CEP_OPT_TIME_POINT && value >= 0.0) b->time_point = value;
  else if (option == CEP_OPT_ORD11_WAVES && (value == 0.0 || value == 1.0 || value == 2.0 || value == 4.0))
    b->o11_waves = (int)value;
  else if (option == CEP_OPT_INTEGRATOR && (value == CEP_INTEGRATOR_RUSH_LARSEN || value == CEP_INTEGRATOR_ADAMS))
    b->integrator = (int)value;
  else return fail(b, CEP_E_ARG, "unknown option or value");
  return 0;
}

static int device_cus()
{
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

// ORd 2011 Rush-Larsen: wavefronts per 64 cells.  Four while four times the wave count still
// fits one wave per SIMD (CUs x 4), else two while that fits, else one, whose waves already
// fill the SIMDs.  CEP_OPT_ORD11_WAVES forces a choice.
static int ord11_waves(const cep_batch *b)
{
  if (b->o11_waves == 1 || b->o11_waves == 2 || b->o11_waves == 4) return b->o11_waves;
  const int simds = 4 * device_cus();
  const int blocks = (b->n + 63) / 64;
  return 4 * blocks <= simds ? 4 : 2 * blocks <= simds ? 2 : 1;
}

// TT06 and the one-wave ORd 2011 kernel: waves per SIMD of the register allocation (1: 512
// registers, no spills; 2: 256 registers, spills to scratch, a second wave hides latency).
// Automatic: 1 while the batch has no more waves than SIMDs (a second wave could not come), else
// 2.  Measured one-beat trials (MI355X): TT06 4,000 cells 194 ms with 1 against 213 with 2,
// 200,000 cells 780 against 631; ORd 2011 100,000 cells 210 against 186.
static int rl_waves_per_eu(const cep_batch *b)
{
  if (b->shape_wpe == 1 || b->shape_wpe == 2) return b->shape_wpe;
  return (b->n + 63) / 64 <= 4 * device_cus() ? 1 : 2;
}

// TT06: threads per workgroup (one cell per thread).  Automatic: 64 (one wave per workgroup, so
// the dispatcher places waves one at a time): 616 ms against 631 with 256 on 200,000 cells, 197
// against 213 on 4,000.
static int tt_block(const cep_batch *b) { return b->shape_block ? b->shape_block : 64; }

// ORd-CiPA whole-beat launches: the order in which the groups of CPB cells are dispatched.
// Automatic: longest beat first when there are more groups than workgroups the device holds at
// once (the dispatcher then refills CUs as groups end, and long groups started last would trail),
// else storage order (every group starts at once; the order changes nothing).
static int bdf_group_order(const cep_batch *b)
{
  if (b->shape_order == CEP_ORDER_STORAGE || b->shape_order == CEP_ORDER_LONGEST_FIRST) return b->shape_order;
  static int resident[2] = {0, 0};
  const int f64 = b->prec == CEP_PREC_FP64;
  if (resident[f64] == 0) {
    int per_cu = 0;
    const hipError_t e = f64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                   &per_cu, ord_bdf_kernel<double, ordbdf::CPB, LAUNCH_BEATS>, 64, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                   &per_cu, ord_bdf_kernel<float, ordbdf::CPB, LAUNCH_BEATS>, 64, 0);
    resident[f64] = (e == hipSuccess && per_cu > 0 ? per_cu : 3) * device_cus();
  }
  const int groups = (b->n + ordbdf::CPB - 1) / ordbdf::CPB;
  return groups > resident[f64] ? CEP_ORDER_LONGEST_FIRST : CEP_ORDER_STORAGE;
}

static double max_bcl(const cep_batch *b)
{
  double m = 0;
  if (b->model == CEP_MODEL_ORD_CIPA) for (auto &c : b->ord_in) m = c.bcl > m ? c.bcl : m;
  else if (b->model == CEP_MODEL_ORD2011) for (auto &c : b->o11_in) m = c.bcl > m ? c.bcl : m;
  else for (auto &c : b->tt_in) m = c.bcl > m ? c.bcl : m;
  return m;
}

static int alloc_trace_buffers(cep_batch *b)
{
  if (b->dr_on && !b->d_dr) {
    if (dalloc(b, &b->d_dr, (size_t)b->pitch * DR_SN)) return CEP_E_HIP;
  }
  if (b->model == CEP_MODEL_ORD2011) {   // no Ca trace; the Adams path may keep a Vm trace
    b->cap = 1;
    const int beat_steps = (int)llround(max_bcl(b) / b->dt);
    if (b->vtrace_on && b->integrator == CEP_INTEGRATOR_ADAMS && beat_steps > b->tcap) {
      if (dalloc(b, &b->d_vtrace, (size_t)b->pitch * beat_steps)) return CEP_E_HIP;
      b->tcap = beat_steps;
    }
    return 0;
  }
  // Ca trace of one beat sampled every 1 ms (dt_write), sized for the longest BCL
  const int print_freq = (int)((1. / b->dt) * 1.0);
  const int beat_steps = (int)(max_bcl(b) / b->dt);
  const int cap = beat_steps / (print_freq > 0 ? print_freq : 1) + 4;
  if (cap > b->cap) {
    int rc = dalloc(b, &b->d_ct, (size_t)b->pitch * cap);
    rc |= dalloc(b, &b->d_cv, (size_t)b->pitch * cap);
    if (rc) return rc;
    b->cap = cap;
  }
  if (b->vtrace_on && beat_steps > b->tcap) {
    if (dalloc(b, &b->d_vtrace, (size_t)b->pitch * beat_steps)) return CEP_E_HIP;
    b->tcap = beat_steps;
  }
  if (b->ts_on && (cap > b->ts_cap || b->nwin > b->ts_nwin)) {
    const int nw = b->nwin > 1 ? b->nwin : 1;
    int rc = dalloc(b, &b->d_ts, (size_t)b->pitch * 2 * cap * 9);
    rc |= dalloc(b, &b->d_tsv, (size_t)b->pitch * nw);
    rc |= dalloc(b, &b->d_tsp, (size_t)b->pitch * nw);
    if (rc) return rc;
    b->ts_cap = cap;
    b->ts_nwin = nw;
  }
  return 0;
}

static int reset_integrators(cep_batch *b)
{
  const size_t P = b->pitch;
  HIPCHK(b, hipMemsetAsync(b->d_status, 0, P * sizeof(int), b->stream));
  HIPCHK(b, hipMemsetAsync(b->d_win, 0xff, P * sizeof(int), b->stream));       // -1: the batch's window
  HIPCHK(b, hipMemsetAsync(b->d_bmd, 0, P * BM_TOTAL_D * sizeof(double), b->stream));
  HIPCHK(b, hipMemsetAsync(b->d_bmi, 0, P * BM_I_N * sizeof(int), b->stream));
  if (b->d_dr) HIPCHK(b, hipMemsetAsync(b->d_dr, 0, P * DR_SN * sizeof(double), b->stream));
  if (b->model == CEP_MODEL_ORD_CIPA) {
    HIPCHK(b, hipMemsetAsync(b->d_si, 0, P * ordbdf::SI_N * sizeof(int), b->stream));   // phase = FRESH
    // the integrator keeps Nordsieck rows above the current order at zero
    HIPCHK(b, hipMemsetAsync(b->d_zn, 0, P * (ordbdf::QMAX + 1) * ordbdf::N * sizeof(double), b->stream));
    HIPCHK(b, hipMemsetAsync(b->d_sd, 0, P * ordbdf::SD_N * sizeof(double), b->stream));
    HIPCHK(b, hipMemsetAsync(b->d_stats, 0, P * 8 * sizeof(long long), b->stream));
  } else if (b->model == CEP_MODEL_ORD2011) {
    HIPCHK(b, hipMemsetAsync(b->d_stats, 0, 2 * P * sizeof(long long), b->stream));
    b->beats_done = 0;
    if (b->integrator == CEP_INTEGRATOR_ADAMS) {
      if (!b->d_aphi) {
        int rc = dalloc(b, &b->d_aphi, P * adams::COLS * ord11::NS);
        rc |= dalloc(b, &b->d_avec, P * 5 * ord11::NS);
        rc |= dalloc(b, &b->d_acoef, P * adams::C_N);
        rc |= dalloc(b, &b->d_asc, P * adams::S_N);
        rc |= dalloc(b, &b->d_asi, P * adams::I_N);
        if (rc) return rc;
      }
      HIPCHK(b, hipMemsetAsync(b->d_aphi, 0, P * adams::COLS * ord11::NS * sizeof(double), b->stream));
      HIPCHK(b, hipMemsetAsync(b->d_avec, 0, P * 5 * ord11::NS * sizeof(double), b->stream));
      HIPCHK(b, hipMemsetAsync(b->d_acoef, 0, P * adams::C_N * sizeof(double), b->stream));
      HIPCHK(b, hipMemsetAsync(b->d_asc, 0, P * adams::S_N * sizeof(double), b->stream));
      HIPCHK(b, hipMemsetAsync(b->d_asi, 0, P * adams::I_N * sizeof(int), b->stream));
    }
  } else {
    HIPCHK(b, hipMemsetAsync(b->d_tcur, 0, P * sizeof(double), b->stream));
    HIPCHK(b, hipMemsetAsync(b->d_iout, 0, P * sizeof(int), b->stream));
  }
  return 0;
}

int cep_init_consts(cep_batch *b)
{
  if (!b) return CEP_E_ARG;
  const size_t P = b->pitch;
  std::vector<double> prm(P * b->np, 0.0), y(P * b->ns, 0.0);
  double p[64];
  for (int i = 0; i < b->n; ++i) {
    if (b->model == CEP_MODEL_ORD_CIPA) {
      host::ord_params(b->ord_in[i], b->stim_scale, p);
      for (int s = 0; s < 49; ++s) y[s * P + i] = host::kOrdInitStates[s];
      y[43 * P + i] = b->ord_in[i].conc;
    } else if (b->model == CEP_MODEL_ORD2011) {
      host::o11_params(b->o11_in[i], b->dutta, b->stim_scale, p);
      for (int s = 0; s < ord11::NS; ++s) y[s * P + i] = host::kO11Init[s];
    } else {
      host::tt_params(b->tt_in[i], b->stim_scale, p);
      const int ct = (int)b->tt_in[i].celltype;
      for (int s = 0; s < tt::NS; ++s) y[s * P + i] = host::kTtInit[ct][s];
    }
    for (int k = 0; k < b->np; ++k) prm[k * P + i] = p[k];
  }
  if (int rc = alloc_trace_buffers(b)) return rc;
  HIPCHK(b, hipMemcpyAsync(b->d_prm, prm.data(), prm.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIPCHK(b, hipMemcpyAsync(b->d_y, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  if (b->model == CEP_MODEL_ORD_CIPA) {
    // the groups of CPB cells by their longest beat, longest first (ties in storage order)
    const int ng = (b->n + ordbdf::CPB - 1) / ordbdf::CPB;
    std::vector<double> glen(ng, 0.0);
    for (int i = 0; i < b->n; ++i) glen[i / ordbdf::CPB] = std::max(glen[i / ordbdf::CPB], b->ord_in[i].bcl);
    b->h_gorder.resize(ng);
    for (int g = 0; g < ng; ++g) b->h_gorder[g] = g;
    std::stable_sort(b->h_gorder.begin(), b->h_gorder.end(), [&](int x, int y) { return glen[x] > glen[y]; });
    if (!b->d_gorder)
      if (int rc = dalloc(b, &b->d_gorder, (size_t)ng)) return rc;
    HIPCHK(b, hipMemcpyAsync(b->d_gorder, b->h_gorder.data(), ng * sizeof(int), hipMemcpyHostToDevice, b->stream));
  }
  if (int rc = reset_integrators(b)) return rc;
  HIPCHK(b, hipStreamSynchronize(b->stream));
  if (const char *cache = getenv("CEP_TUNE_CACHE"))
    if (*cache) (void)cep_tune_lookup(b, cache);
  return 0;
}

int cep_get_constants(const cep_batch *b, int cell, double *out)
{
  if (!b || cell < 0 || cell >= b->n || !out) return CEP_E_ARG;
  if (b->model == CEP_MODEL_ORD_CIPA) host::ord_constants(b->ord_in[cell], b->stim_scale, out);
  else if (b->model == CEP_MODEL_ORD2011) host::o11_constants(b->o11_in[cell], b->dutta, b->stim_scale, out);
  else host::tt_constants(b->tt_in[cell], b->stim_scale, out);
  return 0;
}

int cep_compute_rates(cep_batch *b, double t, const double *states, double *rates, double *algebraic)
{
  if (!b || !states || !rates) return fail(b, CEP_E_ARG, "null array");
  const size_t nsz = (size_t)b->n * b->ns;
  double *dy = nullptr, *df = nullptr, *da = nullptr;
  HIPCHK(b, hipMalloc(&dy, nsz * sizeof(double)));
  HIPCHK(b, hipMalloc(&df, nsz * sizeof(double)));
  if (algebraic) {
    HIPCHK(b, hipMalloc(&da, (size_t)b->n * b->nalg * sizeof(double)));
    HIPCHK(b, hipMemsetAsync(da, 0, (size_t)b->n * b->nalg * sizeof(double), b->stream));
  }
  HIPCHK(b, hipMemcpyAsync(dy, states, nsz * sizeof(double), hipMemcpyHostToDevice, b->stream));
  // the rates kernels are compiled for 64-thread workgroups (__launch_bounds__(64))
  constexpr int kRatesBlock = 64;
  const int blocks = (b->n + kRatesBlock - 1) / kRatesBlock;
  if (b->model == CEP_MODEL_ORD_CIPA) {
    if (b->prec == CEP_PREC_FP64)
      hipLaunchKernelGGL(ord_rates_kernel<double>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
    else
      hipLaunchKernelGGL(ord_rates_kernel<float>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
  } else if (b->model == CEP_MODEL_ORD2011) {
    if (b->prec == CEP_PREC_FP64)
      hipLaunchKernelGGL(ord11_rates_kernel<double>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
    else
      hipLaunchKernelGGL(ord11_rates_kernel<float>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
  } else {
    if (b->prec == CEP_PREC_FP64)
      hipLaunchKernelGGL(tt_rates_kernel<double>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
    else
      hipLaunchKernelGGL(tt_rates_kernel<float>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
  }
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipMemcpyAsync(rates, df, nsz * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  if (algebraic) HIPCHK(b, hipMemcpyAsync(algebraic, da, (size_t)b->n * b->nalg * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(b, hipStreamSynchronize(b->stream));
  (void)hipFree(dy);
  (void)hipFree(df);
  if (da) (void)hipFree(da);
  return 0;
}

static int launch_ord11_adams(cep_batch *b, int nbeats, int nout_fixed)
{
  // the Adams workspace is allocated by cep_init_consts when the integrator is Adams; a switch
  // of integrator after it leaves the workspace unallocated until the next cep_init_consts
  if (!b->d_aphi || !b->d_avec || !b->d_acoef || !b->d_asc || !b->d_asi)
    return fail(b, CEP_E_STATE, "the Adams integrator was selected after cep_init_consts: call cep_init_consts again");
  O11Buffers B;
  B.ncell = b->n; B.pitch = b->pitch; B.nbeats_total = b->nbeats_total; B.nwin = b->nwin;
  B.y = b->d_y; B.prm = b->d_prm; B.bmd = b->d_bmd; B.bmi = b->d_bmi; B.status = b->d_status;
  B.stats = b->d_stats; B.time_point = b->time_point; B.max_dt = b->dt;
  adams::Buffers A;
  A.ncell = b->n; A.pitch = b->pitch; A.phi = b->d_aphi; A.vec = b->d_avec; A.coef = b->d_acoef;
  A.sc = b->d_asc; A.si = b->d_asi; A.stats = b->d_stats;
  double *vt = b->vtrace_on ? b->d_vtrace : nullptr;
  const int blocks = (b->n + 63) / 64;
  if (b->prec == CEP_PREC_FP64)
    hipLaunchKernelGGL(ord11_adams_kernel<double>, dim3(blocks), dim3(64), 0, b->stream, B, A, nbeats, nout_fixed,
                       b->dt, b->rtol, b->atol, vt, vt ? b->tcap : 0);
  else
    hipLaunchKernelGGL(ord11_adams_kernel<float>, dim3(blocks), dim3(64), 0, b->stream, B, A, nbeats, nout_fixed,
                       b->dt, b->rtol, b->atol, vt, vt ? b->tcap : 0);
  return 0;
}

// the ORd-CiPA BDF launch: nbeats beats of each cell's BCL, or nout_rel further output points,
// or up to the absolute output point iout_abs (exactly one of the three is positive)
static void launch_ord_bdf(cep_batch *b, const RunCfg &rc, int nbeats, int nout_rel, int iout_abs)
{
  Buffers B;
  B.ncell = b->n; B.pitch = b->pitch; B.y = b->d_y; B.zn = b->d_zn; B.ewt = b->d_ewt;
  B.acor = b->d_acor; B.ft = b->d_ft; B.jd = b->d_jd; B.fac = b->d_fac; B.sd = b->d_sd;
  B.si = b->d_si; B.prm = b->d_prm; B.bmd = b->d_bmd; B.bmi = b->d_bmi; B.ct = b->d_ct;
  B.cv = b->d_cv; B.vtrace = b->vtrace_on ? b->d_vtrace : nullptr; B.tcap = b->tcap;
  B.stats = b->d_stats; B.status = b->d_status; B.dt = b->dt; B.rtol = b->rtol; B.atol = b->atol;
  B.uround_fd = b->prec == CEP_PREC_FP64 ? ordbdf::UROUND : 1.1920928955078125e-07;
  B.maxord = b->maxord; B.max_steps = 500; B.cap = b->cap;
  B.win = b->d_win;
  B.ts = TsView{};
  if (b->ts_on && b->d_ts) {
    const size_t sl = (size_t)b->pitch * 2 * b->ts_cap;   // one field, both slots
    B.ts.t = b->d_ts; B.ts.vm = b->d_ts + sl; B.ts.ca = b->d_ts + 2 * sl; B.ts.ires = b->d_ts + 3 * sl;
    B.ts.rp = b->d_tsp; B.ts.rv = b->d_tsv; B.ts.cap = b->ts_cap; B.ts.nwin = b->ts_nwin;
  }
  // safety net only: a step takes one pass when Newton converges at once, a few more on
  // failures; cells still running after this many passes are flagged CEP_ST_MAX_STEPS
  const long long nout = nbeats > 0 ? (long long)nbeats * (long long)(max_bcl(b) / b->dt)
                                    : nout_rel > 0 ? nout_rel : iout_abs;
  const long long max_passes = 64LL * nout + 100000;
  constexpr int CPW = ordbdf::CPB;
  const int blocks = (b->n + CPW - 1) / CPW;
  const dim3 g(blocks), w(64);
  const bool f64 = b->prec == CEP_PREC_FP64;
  // the driver rules' state is dynamic LDS, allocated only when the rules are on
  const unsigned dyn = rc.dr ? CPW * (unsigned)(sizeof(DriverRules) + 3 * sizeof(double)) : 0u;
  B.gorder = nullptr;
  if (nbeats > 0 && bdf_group_order(b) == CEP_ORDER_LONGEST_FIRST) {
    B.gorder = b->d_gorder;
    if (f64) hipLaunchKernelGGL((ord_bdf_kernel<double, CPW, LAUNCH_BEATS_ORDERED>), g, w, dyn, b->stream, B, rc, nbeats, max_passes);
    else hipLaunchKernelGGL((ord_bdf_kernel<float, CPW, LAUNCH_BEATS_ORDERED>), g, w, dyn, b->stream, B, rc, nbeats, max_passes);
  } else if (nbeats > 0) {
    if (f64) hipLaunchKernelGGL((ord_bdf_kernel<double, CPW, LAUNCH_BEATS>), g, w, dyn, b->stream, B, rc, nbeats, max_passes);
    else hipLaunchKernelGGL((ord_bdf_kernel<float, CPW, LAUNCH_BEATS>), g, w, dyn, b->stream, B, rc, nbeats, max_passes);
  } else if (nout_rel > 0) {
    if (f64) hipLaunchKernelGGL((ord_bdf_kernel<double, CPW, LAUNCH_OUTPUTS>), g, w, dyn, b->stream, B, rc, nout_rel, max_passes);
    else hipLaunchKernelGGL((ord_bdf_kernel<float, CPW, LAUNCH_OUTPUTS>), g, w, dyn, b->stream, B, rc, nout_rel, max_passes);
  } else {
    if (f64) hipLaunchKernelGGL((ord_bdf_kernel<double, CPW, LAUNCH_TO_OUTPUT>), g, w, dyn, b->stream, B, rc, iout_abs, max_passes);
    else hipLaunchKernelGGL((ord_bdf_kernel<float, CPW, LAUNCH_TO_OUTPUT>), g, w, dyn, b->stream, B, rc, iout_abs, max_passes);
  }
}

static RunCfg run_cfg(const cep_batch *b)
{
  RunCfg rc{b->dt, b->nbeats_total, b->nwin, (int)((1. / b->dt) * 1.0), b->t_first > 0.0 ? b->t_first : b->dt,
            b->v_apd90, b->v_apd50, b->dr_on ? b->d_dr : nullptr};
  if (rc.print_freq < 1) rc.print_freq = 1;
  return rc;
}

int cep_run_outputs(cep_batch *b, int nout)
{
  if (!b || nout < 1) return fail(b, CEP_E_ARG, "nout");
  if (b->cap == 0) return fail(b, CEP_E_STATE, "cep_init_consts not called");
  if (b->model == CEP_MODEL_ORD_CIPA) {
    if (int rc = alloc_trace_buffers(b)) return rc;
    HIPCHK(b, hipEventRecord(b->ev0, b->stream));
    launch_ord_bdf(b, run_cfg(b), 0, nout, 0);
    HIPCHK(b, hipGetLastError());
    HIPCHK(b, hipEventRecord(b->ev1, b->stream));
    b->pending = true;
    return cep_sync(b);
  }
  if (b->model != CEP_MODEL_ORD2011 || b->integrator != CEP_INTEGRATOR_ADAMS)
    return fail(b, CEP_E_STATE, "output-step runs are the BDF's (ORd-CiPA) and the Adams integrator's");
  if (!b->d_aphi)
    return fail(b, CEP_E_STATE, "the Adams integrator was selected after cep_init_consts: call cep_init_consts again");
  if (int rc = alloc_trace_buffers(b)) return rc;
  HIPCHK(b, hipEventRecord(b->ev0, b->stream));
  if (int rc = launch_ord11_adams(b, 0, nout)) return rc;
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipEventRecord(b->ev1, b->stream));
  b->pending = true;
  return cep_sync(b);
}

int cep_integrate_bdf(cep_batch *b, double tout)
{
  if (!b) return CEP_E_ARG;
  if (b->model != CEP_MODEL_ORD_CIPA) return fail(b, CEP_E_ARG, "the BDF integrates ORd-CiPA cells");
  if (b->cap == 0) return fail(b, CEP_E_STATE, "cep_init_consts not called");
  const double t1 = b->t_first > 0.0 ? b->t_first : b->dt;
  const double k = (tout - t1) / b->dt;
  const long long kr = llround(k);
  if (!(tout >= t1) || fabs(k - (double)kr) > 1e-6 || kr + 1 > 2000000000LL)
    return fail(b, CEP_E_ARG, "tout must be an output point t_first + k dt");
  if (int rc = alloc_trace_buffers(b)) return rc;
  HIPCHK(b, hipEventRecord(b->ev0, b->stream));
  launch_ord_bdf(b, run_cfg(b), 0, 0, (int)(kr + 1));
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipEventRecord(b->ev1, b->stream));
  b->pending = true;
  return cep_sync(b);
}

int cep_run_beats_async(cep_batch *b, int nbeats)
{
  if (!b || nbeats < 1) return fail(b, CEP_E_ARG, "nbeats");
  if (b->cap == 0) return fail(b, CEP_E_STATE, "cep_init_consts not called");
  if (b->model == CEP_MODEL_ORD2011 && b->integrator == CEP_INTEGRATOR_ADAMS && !b->d_aphi)
    return fail(b, CEP_E_STATE, "the Adams integrator was selected after cep_init_consts: call cep_init_consts again");
  if (int rc = alloc_trace_buffers(b)) return rc;
  if (b->dr_on && (b->model == CEP_MODEL_ORD_CIPA || b->model == CEP_MODEL_TT06) && b->nbeats_total < 2)
    return fail(b, CEP_E_ARG, "the driver rules read beats n-2 and n-1: nbeats_total >= 2");
  const RunCfg rc = run_cfg(b);
  HIPCHK(b, hipEventRecord(b->ev0, b->stream));
  if (b->model == CEP_MODEL_ORD_CIPA) {
    launch_ord_bdf(b, rc, nbeats, 0, 0);
  } else if (b->model == CEP_MODEL_ORD2011 && b->integrator == CEP_INTEGRATOR_ADAMS) {
    if (int r = launch_ord11_adams(b, nbeats, 0)) return r;
    b->beats_done += nbeats;
  } else if (b->model == CEP_MODEL_ORD2011) {
    O11Buffers B;
    B.ncell = b->n; B.pitch = b->pitch; B.nbeats_total = b->nbeats_total; B.nwin = b->nwin;
    B.y = b->d_y; B.prm = b->d_prm; B.bmd = b->d_bmd; B.bmi = b->d_bmi; B.status = b->d_status;
    B.stats = b->d_stats; B.time_point = b->time_point; B.max_dt = b->dt;
    // guard only: the smallest step the rule produces is 0.0005 ms
    const long long max_steps = (long long)nbeats * (long long)(max_bcl(b) / 0.0005) + 1000;
    const int blocks = (b->n + 63) / 64;
    const int waves = ord11_waves(b);
    if (waves == 4) {
      if (b->prec == CEP_PREC_FP64)
        hipLaunchKernelGGL(ord11_rl4_kernel<double>, dim3(blocks), dim3(256), 0, b->stream, B, b->beats_done + nbeats, max_steps);
      else
        hipLaunchKernelGGL(ord11_rl4_kernel<float>, dim3(blocks), dim3(256), 0, b->stream, B, b->beats_done + nbeats, max_steps);
    } else if (waves == 2) {
      if (b->prec == CEP_PREC_FP64)
        hipLaunchKernelGGL(ord11_rl2_kernel<double>, dim3(blocks), dim3(128), 0, b->stream, B, b->beats_done + nbeats, max_steps);
      else
        hipLaunchKernelGGL(ord11_rl2_kernel<float>, dim3(blocks), dim3(128), 0, b->stream, B, b->beats_done + nbeats, max_steps);
    } else if (rl_waves_per_eu(b) == 1) {
      if (b->prec == CEP_PREC_FP64)
        hipLaunchKernelGGL((ord11_rl_kernel<double, 1>), dim3(blocks), dim3(64), 0, b->stream, B, b->beats_done + nbeats, max_steps);
      else
        hipLaunchKernelGGL((ord11_rl_kernel<float, 1>), dim3(blocks), dim3(64), 0, b->stream, B, b->beats_done + nbeats, max_steps);
    } else if (b->prec == CEP_PREC_FP64) {
      hipLaunchKernelGGL((ord11_rl_kernel<double, 2>), dim3(blocks), dim3(64), 0, b->stream, B, b->beats_done + nbeats, max_steps);
    } else {
      hipLaunchKernelGGL((ord11_rl_kernel<float, 2>), dim3(blocks), dim3(64), 0, b->stream, B, b->beats_done + nbeats, max_steps);
    }
    b->beats_done += nbeats;
  } else {
    TtBuffers B;
    B.ncell = b->n; B.pitch = b->pitch; B.cap = b->cap; B.tcap = b->tcap; B.y = b->d_y;
    B.prm = b->d_prm; B.bmd = b->d_bmd; B.ct = b->d_ct; B.cv = b->d_cv; B.tcur = b->d_tcur;
    B.vtrace = b->vtrace_on ? b->d_vtrace : nullptr; B.bmi = b->d_bmi; B.iout = b->d_iout;
    B.status = b->d_status; B.win = b->d_win;
    const int tb = tt_block(b);
    const int blocks = (b->n + tb - 1) / tb;
    const bool one = rl_waves_per_eu(b) == 1;
    if (b->prec == CEP_PREC_FP64) {
      if (one) hipLaunchKernelGGL((tt_rl_kernel<double, 1>), dim3(blocks), dim3(tb), 0, b->stream, B, rc, nbeats);
      else hipLaunchKernelGGL((tt_rl_kernel<double, 2>), dim3(blocks), dim3(tb), 0, b->stream, B, rc, nbeats);
    } else {
      if (one) hipLaunchKernelGGL((tt_rl_kernel<float, 1>), dim3(blocks), dim3(tb), 0, b->stream, B, rc, nbeats);
      else hipLaunchKernelGGL((tt_rl_kernel<float, 2>), dim3(blocks), dim3(tb), 0, b->stream, B, rc, nbeats);
    }
  }
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipEventRecord(b->ev1, b->stream));
  b->pending = true;
  return 0;
}

int cep_sync(cep_batch *b)
{
  if (!b) return CEP_E_ARG;
  HIPCHK(b, hipStreamSynchronize(b->stream));
  if (b->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) { b->ktime_total += ms; b->klaunches++; }
    b->pending = false;
  }
  return 0;
}

int cep_query(cep_batch *b)
{
  if (!b) return CEP_E_ARG;
  if (!b->pending) return 0;
  const hipError_t e = hipEventQuery(b->ev1);
  if (e == hipErrorNotReady) return 1;
  if (e != hipSuccess) return fail(b, CEP_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
  return cep_sync(b) == 0 ? 0 : CEP_E_HIP;
}

long long cep_batch_device_bytes(int model, int ncells, double max_bcl, double dt, int vm_trace)
{
  if ((model != CEP_MODEL_ORD_CIPA && model != CEP_MODEL_TT06 && model != CEP_MODEL_ORD2011) || ncells <= 0 ||
      !(dt > 0) || !(max_bcl > 0))
    return CEP_E_ARG;
  if (model == CEP_MODEL_ORD2011) {
    const long long P = (ncells + 63) / 64 * 64;
    return (P * (ord11::NS + ord11::NP + BM_TOTAL_D + 1)) * 8 + P * (BM_I_N + 1) * 4 + 6 * 64;
  }
  const long long P = (ncells + 63) / 64 * 64, pad = 64;
  const bool ord = model == CEP_MODEL_ORD_CIPA;
  const long long ns = ord ? ord::NS : tt::NS, np = ord ? ord::NP : tt::NP;
  long long d = P * ns + P * np + P * BM_TOTAL_D;            // doubles
  long long i4 = P * BM_I_N + P;                              // ints
  if (ord) {
    d += P * ((ordbdf::QMAX + 1) * ordbdf::N + 3 * ordbdf::N + ordbdf::NJC * ordbdf::N + ordbdf::NFAC + ordbdf::SD_N);
    i4 += P * ordbdf::SI_N;
    d += P * 8;                                               // stats (long long)
  } else {
    d += P;
    i4 += P;
  }
  const int print_freq = (int)((1. / dt) * 1.0);
  const long long beat_steps = (long long)(max_bcl / dt);
  const long long cap = beat_steps / (print_freq > 0 ? print_freq : 1) + 4;
  d += 2 * P * cap;
  if (vm_trace) d += P * beat_steps;
  d += P * DR_SN;                                             // driver-rule state, if enabled
  const int nalloc = ord ? 15 : 11;
  return d * 8 + i4 * 4 + nalloc * pad;
}

int cep_run_beats(cep_batch *b, int nbeats)
{
  int rc = cep_run_beats_async(b, nbeats);
  if (rc) return rc;
  return cep_sync(b);
}

static int get_soa(const cep_batch *b, const double *d, int width, double *out)
{
  std::vector<double> tmp((size_t)b->pitch * width);
  if (hipMemcpyAsync(tmp.data(), d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return CEP_E_HIP;
  for (int i = 0; i < b->n; ++i)
    for (int k = 0; k < width; ++k) out[(size_t)i * width + k] = tmp[(size_t)k * b->pitch + i];
  return 0;
}

int cep_get_states(const cep_batch *b, double *out)
{
  if (!b || !out) return CEP_E_ARG;
  return get_soa(b, b->d_y, b->ns, out);
}

int cep_set_states(cep_batch *b, const double *in)
{
  if (!b || !in) return CEP_E_ARG;
  std::vector<double> tmp((size_t)b->pitch * b->ns, 0.0);
  for (int i = 0; i < b->n; ++i)
    for (int k = 0; k < b->ns; ++k) tmp[(size_t)k * b->pitch + i] = in[(size_t)i * b->ns + k];
  HIPCHK(b, hipMemcpyAsync(b->d_y, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  if (int rc = reset_integrators(b)) return rc;
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int cep_get_biomarkers(const cep_batch *b, double *out)
{
  if (!b || !out) return CEP_E_ARG;
  std::vector<double> d((size_t)b->pitch * BM_TOTAL_D);
  std::vector<int> iv((size_t)b->pitch * BM_I_N);
  if (hipMemcpyAsync(d.data(), b->d_bmd, d.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipMemcpyAsync(iv.data(), b->d_bmi, iv.size() * sizeof(int), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return CEP_E_HIP;
  for (int i = 0; i < b->n; ++i) {
    if (b->model == CEP_MODEL_ORD2011) {
      for (int k = 0; k < CEP_BM11_STEEPEST_PACE; ++k) out[(size_t)i * CEP_NBM + k] = d[(size_t)(O11_BEST + k) * b->pitch + i];
      out[(size_t)i * CEP_NBM + CEP_BM11_STEEPEST_PACE] = iv[(size_t)O11_I_STEEP * b->pitch + i];
      continue;
    }
    for (int k = 0; k < BM_RES_N; ++k) out[(size_t)i * CEP_NBM + k] = d[(size_t)(BM_D_N + BM_RES_N + k) * b->pitch + i];
    out[(size_t)i * CEP_NBM + CEP_BM_STEEPEST_PACE] = iv[(size_t)1 * b->pitch + i];
  }
  return 0;
}

int cep_get_status(const cep_batch *b, int *status)
{
  if (!b || !status) return CEP_E_ARG;
  if (hipMemcpyAsync(status, b->d_status, (size_t)b->n * sizeof(int), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return CEP_E_HIP;
  return 0;
}

int cep_get_stats(const cep_batch *b, long long *out)
{
  if (!b || !out) return CEP_E_ARG;
  if (!b->d_stats) { memset(out, 0, sizeof(long long) * 8 * b->n); return 0; }
  if (b->model == CEP_MODEL_ORD2011) {   // steps (one right-hand side each) in slots 0 and 1
    std::vector<long long> tmp((size_t)b->pitch * 2);
    if (hipMemcpyAsync(tmp.data(), b->d_stats, tmp.size() * sizeof(long long), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return CEP_E_HIP;
    memset(out, 0, sizeof(long long) * 8 * b->n);
    const bool adams_path = b->integrator == CEP_INTEGRATOR_ADAMS;
    for (int i = 0; i < b->n; ++i) {
      out[(size_t)i * 8] = tmp[i];
      out[(size_t)i * 8 + 1] = adams_path ? tmp[b->pitch + i] : tmp[i];
    }
    return 0;
  }
  std::vector<long long> tmp((size_t)b->pitch * 8);
  if (hipMemcpyAsync(tmp.data(), b->d_stats, tmp.size() * sizeof(long long), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return CEP_E_HIP;
  for (int i = 0; i < b->n; ++i)
    for (int k = 0; k < 8; ++k) out[(size_t)i * 8 + k] = tmp[(size_t)k * b->pitch + i];
  return 0;
}

int cep_enable_vm_trace(cep_batch *b, int enable)
{
  if (!b) return CEP_E_ARG;
  b->vtrace_on = enable != 0;
  return 0;
}

int cep_get_vm_trace(const cep_batch *b, double *out, int *nsamp)
{
  if (!b || !nsamp) return CEP_E_ARG;
  if (!b->vtrace_on || !b->d_vtrace) return fail(const_cast<cep_batch *>(b), CEP_E_STATE, "Vm trace not enabled");
  *nsamp = b->tcap;
  if (!out) return 0;   // size query
  return get_soa(b, b->d_vtrace, b->tcap, out);
}

// one cell's column of a SoA block: rows [0, n) of width 1 at row pitch b->pitch
static int get_column(const cep_batch *b, const void *d, size_t elem, int cell, int n, void *out)
{
  if (n <= 0) return 0;
  if (hipMemcpy2DAsync(out, elem, (const char *)d + (size_t)cell * elem, (size_t)b->pitch * elem, elem, n,
                       hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return CEP_E_HIP;
  return 0;
}

int cep_get_time_series(const cep_batch *b, int cell, int max_n, double *t, double *vm, double *cai, double *ires,
                        int *n, int *pace)
{
  if (!b || !n || cell < 0 || cell >= b->n || max_n < 0) return CEP_E_ARG;
  cep_batch *bm = const_cast<cep_batch *>(b);
  if (!b->ts_on || !b->d_ts) return fail(bm, CEP_E_STATE, "time series not enabled (CEP_OPT_TIME_SERIES)");
  int iv[BM_I_N];
  if (int rc = get_column(b, b->d_bmi, sizeof(int), cell, BM_I_N, iv)) return rc;
  const int slot = iv[5], cnt = iv[6] < b->ts_cap ? iv[6] : b->ts_cap;
  *n = cnt;
  if (pace) *pace = iv[1];
  if (!t && !vm && !cai && !ires) return 0;   // size query
  const int m = cnt < max_n ? cnt : max_n;
  const size_t sl = (size_t)b->pitch * 2 * b->ts_cap, off = (size_t)slot * b->ts_cap * b->pitch;
  if (t) if (int rc = get_column(b, b->d_ts + off, 8, cell, m, t)) return rc;
  if (vm) if (int rc = get_column(b, b->d_ts + sl + off, 8, cell, m, vm)) return rc;
  if (cai) if (int rc = get_column(b, b->d_ts + 2 * sl + off, 8, cell, m, cai)) return rc;
  if (ires) {
    std::vector<double> col((size_t)m);
    for (int j = 0; j < 6; ++j) {
      const size_t oj = 3 * sl + ((size_t)(slot * 6 + j) * b->ts_cap) * b->pitch;
      if (int rc = get_column(b, b->d_ts + oj, 8, cell, m, col.data())) return rc;
      for (int k = 0; k < m; ++k) ires[(size_t)k * 6 + j] = col[k];
    }
  }
  return 0;
}

int cep_get_dvmdt_repol_history(const cep_batch *b, int cell, int max_n, int *paces, double *values, int *n)
{
  if (!b || !n || cell < 0 || cell >= b->n || max_n < 0) return CEP_E_ARG;
  cep_batch *bm = const_cast<cep_batch *>(b);
  if (!b->ts_on || !b->d_ts) return fail(bm, CEP_E_STATE, "time series not enabled (CEP_OPT_TIME_SERIES)");
  int iv[BM_I_N];
  if (int rc = get_column(b, b->d_bmi, sizeof(int), cell, BM_I_N, iv)) return rc;
  const int cnt = iv[7] < b->ts_nwin ? iv[7] : b->ts_nwin;
  *n = cnt;
  const int m = cnt < max_n ? cnt : max_n;
  if (paces) if (int rc = get_column(b, b->d_tsp, sizeof(int), cell, m, paces)) return rc;
  if (values) if (int rc = get_column(b, b->d_tsv, 8, cell, m, values)) return rc;
  return 0;
}

int cep_get_driver_rules(const cep_batch *b, double *out)
{
  if (!b || !out) return CEP_E_ARG;
  if (!b->d_dr) return fail(const_cast<cep_batch *>(b), CEP_E_STATE, "CEP_OPT_DRIVER_RULES was not on");
  const size_t P = b->pitch;
  std::vector<double> h(P * DR_SN);
  HIPCHK(const_cast<cep_batch *>(b), hipMemcpyAsync(h.data(), b->d_dr, h.size() * sizeof(double),
                                                  hipMemcpyDeviceToHost, b->stream));
  HIPCHK(const_cast<cep_batch *>(b), hipStreamSynchronize(b->stream));
  for (int i = 0; i < b->n; ++i) {
    double st[DR_SN];
    for (int k = 0; k < DR_SN; ++k) st[k] = h[k * P + i];
    if (b->model == CEP_MODEL_ORD_CIPA) {
      DriverRules r;
      memcpy(&r, st, sizeof(r));
      r.record(out + (size_t)i * CEP_NDR);
    } else {
      TnRules r;
      memcpy(&r, st, sizeof(r));
      r.record(out + (size_t)i * CEP_NDR_TT06);
    }
  }
  return 0;
}

}  // extern "C"

// qinward of each cell against its control cell (drugs_herg.cpp:309-315), from the AUC windows
__global__ void __launch_bounds__(64) qinward_kernel(int n, int pitch, const DriverRules *__restrict__ dummy,
                                                    const double *__restrict__ dr, const int *__restrict__ ctl,
                                                    double *__restrict__ out)
{
  (void)dummy;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int k = ctl[c];
  if (k < 0 || k >= n) { out[c] = __builtin_nan(""); return; }
  auto row = [&](int slot, int cell) { return dr[(size_t)slot * pitch + cell]; };
  constexpr int N0 = offsetof(DriverRules, dh_n0) / 8, N1 = offsetof(DriverRules, dh_n1) / 8;
  constexpr int C0 = offsetof(DriverRules, dh_c0) / 8, C1 = offsetof(DriverRules, dh_c1) / 8;
  const double inal_d = row(N1, c) - row(N0, c), ical_d = row(C1, c) - row(C0, c);
  const double inal_c = row(N1, k) - row(N0, k), ical_c = row(C1, k) - row(C0, k);
  out[c] = ((inal_d / inal_c) + (ical_d / ical_c)) * 0.5;
}

extern "C" {

int cep_get_qinward(cep_batch *b, const int *control, double *out)
{
  if (!b || !control || !out) return CEP_E_ARG;
  if (b->model != CEP_MODEL_ORD_CIPA) return fail(b, CEP_E_ARG, "qinward is an ORd-CiPA output");
  if (!b->d_dr) return fail(b, CEP_E_STATE, "CEP_OPT_DRIVER_RULES was not on");
  for (int i = 0; i < b->n; ++i)
    if (control[i] >= b->n) return fail(b, CEP_E_ARG, "control cell index");
  if (!b->d_qin) {
    if (dalloc(b, &b->d_qin, (size_t)b->pitch) || dalloc(b, &b->d_qctl, (size_t)b->pitch)) return CEP_E_HIP;
  }
  HIPCHK(b, hipMemcpyAsync(b->d_qctl, control, b->n * sizeof(int), hipMemcpyHostToDevice, b->stream));
  const int blocks = (b->n + 63) / 64;
  hipLaunchKernelGGL(qinward_kernel, dim3(blocks), dim3(64), 0, b->stream, b->n, b->pitch, nullptr, b->d_dr,
                     b->d_qctl, b->d_qin);
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipMemcpyAsync(out, b->d_qin, b->n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int cep_kernel_time(cep_batch *b, double *avg_ms, double *total_ms)
{
  if (!b) return CEP_E_ARG;
  const int n = b->klaunches;
  if (avg_ms) *avg_ms = n ? b->ktime_total / n : 0.0;
  if (total_ms) *total_ms = b->ktime_total;
  b->ktime_total = 0.0;
  b->klaunches = 0;
  return n;
}

}  // extern "C"

// ---- cells restarted and parked one by one (a resident batch whose cells serve jobs)
//
// A restart gives each listed cell its packed row (states, then parameters), clears every
// per-cell row of the integrator and the biomarker machine, and sets its status and its own
// first window beat.  The integrator then starts as for a fresh batch (phase FRESH: CVodeInit
// at t = 0 from the new states; TT06: output index 0), so a restarted cell computes exactly
// what the same cell computes in a batch of its own.
struct RestartArgs {
  int m, pitch, ns, np, win;
  const int *idx;          // [m] cells
  const double *rows;      // [m][ns + np]
  double *y, *prm, *bmd, *dr;
  int *bmi, *status, *winv;
  double *zn, *sd;         // ORd-CiPA BDF (zn in blocks of ordbdf::CPB cells), else null
  int *si;
  long long *stats;
  double *tcur;            // TT06, else null
  int *iout;
};

__global__ void __launch_bounds__(64) restart_cells_kernel(RestartArgs A)
{
  const int k = blockIdx.x;
  const int c = A.idx[k];
  const size_t P = A.pitch;
  const double *r = A.rows + (size_t)k * (A.ns + A.np);
  for (int j = threadIdx.x; j < A.ns; j += 64) A.y[j * P + c] = r[j];
  for (int j = threadIdx.x; j < A.np; j += 64) A.prm[j * P + c] = r[A.ns + j];
  for (int j = threadIdx.x; j < BM_TOTAL_D; j += 64) A.bmd[j * P + c] = 0.0;
  for (int j = threadIdx.x; j < BM_I_N; j += 64) A.bmi[j * P + c] = 0;
  if (A.dr)
    for (int j = threadIdx.x; j < DR_SN; j += 64) A.dr[j * P + c] = 0.0;
  if (A.zn) {
    using namespace ordbdf;
    double *blk = A.zn + (size_t)(c / CPB) * CPB * ((QMAX + 1) * N) + c % CPB;
    for (int j = threadIdx.x; j < (QMAX + 1) * N; j += 64) blk[(size_t)j * CPB] = 0.0;
    for (int j = threadIdx.x; j < SD_N; j += 64) A.sd[j * P + c] = 0.0;
    for (int j = threadIdx.x; j < SI_N; j += 64) A.si[j * P + c] = 0;    // phase FRESH
    for (int j = threadIdx.x; j < 8; j += 64) A.stats[j * P + c] = 0;
  }
  if (threadIdx.x == 0) {
    if (A.tcur) { A.tcur[c] = 0.0; A.iout[c] = 0; }
    A.status[c] = ST_OK;
    A.winv[c] = A.win;
  }
}

__global__ void __launch_bounds__(64) park_cells_kernel(int m, const int *__restrict__ idx, int *__restrict__ status)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) status[idx[k]] = ST_IDLE;
}

// the listed cells, checked (in range, no repeats) and copied to the staging index buffer
static int stage_cells(cep_batch *b, const int *idx, int m)
{
  if (!idx || m < 1 || m > b->n) return fail(b, CEP_E_ARG, "cell list");
  std::vector<char> seen(b->n, 0);
  for (int k = 0; k < m; ++k) {
    if (idx[k] < 0 || idx[k] >= b->n) return fail(b, CEP_E_ARG, "cell index out of range");
    if (seen[idx[k]]++) return fail(b, CEP_E_ARG, "cell listed twice");
  }
  if (!b->d_stage_idx) {
    // sized once for the whole batch: later restarts and parks reuse them
    if (dalloc(b, &b->d_stage_idx, (size_t)b->n) ||
        dalloc(b, &b->d_stage, (size_t)b->n * (size_t)(b->ns + b->np)))
      return CEP_E_HIP;
  }
  b->h_stage_idx.assign(idx, idx + m);
  HIPCHK(b, hipMemcpyAsync(b->d_stage_idx, b->h_stage_idx.data(), m * sizeof(int), hipMemcpyHostToDevice, b->stream));
  return 0;
}

extern "C" {

int cep_restart_cells(cep_batch *b, const int *idx, int m, int nbeats_total, int nwin)
{
  if (!b) return CEP_E_ARG;
  if (b->model != CEP_MODEL_ORD_CIPA && b->model != CEP_MODEL_TT06)
    return fail(b, CEP_E_ARG, "cells are restarted one by one in ORd-CiPA and TT06 batches");
  if (b->cap == 0) return fail(b, CEP_E_STATE, "cep_init_consts not called");
  if (b->pending) return fail(b, CEP_E_STATE, "a run is in flight: cep_sync or cep_query first");
  if (b->dr_on) return fail(b, CEP_E_STATE, "the driver rules count beats for the whole batch");
  if (nbeats_total < 1 || nwin < 1) return fail(b, CEP_E_ARG, "nbeats_total and nwin");
  if (int rc = stage_cells(b, idx, m)) return rc;
  const size_t row = (size_t)(b->ns + b->np);
  b->h_stage.assign((size_t)m * row, 0.0);
  // the Ca trace was sized at cep_init_consts for the batch's longest beat
  const int print_freq = std::max(1, (int)((1. / b->dt) * 1.0));
  double p[64];
  for (int k = 0; k < m; ++k) {
    const int i = idx[k];
    double *y = &b->h_stage[(size_t)k * row], *q = y + b->ns;
    double bcl;
    if (b->model == CEP_MODEL_ORD_CIPA) {
      host::ord_params(b->ord_in[i], b->stim_scale, p);
      for (int s = 0; s < 49; ++s) y[s] = host::kOrdInitStates[s];
      y[43] = b->ord_in[i].conc;
      bcl = b->ord_in[i].bcl;
    } else {
      host::tt_params(b->tt_in[i], b->stim_scale, p);
      const int ct = (int)b->tt_in[i].celltype;
      for (int s = 0; s < tt::NS; ++s) y[s] = host::kTtInit[ct][s];
      bcl = b->tt_in[i].bcl;
    }
    if ((int)(bcl / b->dt) / print_freq + 4 > b->cap)
      return fail(b, CEP_E_ARG, "a restarted cell's BCL is longer than the batch's at cep_init_consts");
    for (int j = 0; j < b->np; ++j) q[j] = p[j];
  }
  HIPCHK(b, hipMemcpyAsync(b->d_stage, b->h_stage.data(), b->h_stage.size() * sizeof(double), hipMemcpyHostToDevice,
                           b->stream));
  RestartArgs A{};
  A.m = m; A.pitch = b->pitch; A.ns = b->ns; A.np = b->np;
  A.win = nbeats_total - std::min(nwin, nbeats_total);
  A.idx = b->d_stage_idx; A.rows = b->d_stage;
  A.y = b->d_y; A.prm = b->d_prm; A.bmd = b->d_bmd; A.dr = b->d_dr; A.bmi = b->d_bmi;
  A.status = b->d_status; A.winv = b->d_win;
  if (b->model == CEP_MODEL_ORD_CIPA) {
    A.zn = b->d_zn; A.sd = b->d_sd; A.si = b->d_si; A.stats = b->d_stats;
  } else {
    A.tcur = b->d_tcur; A.iout = b->d_iout;
  }
  hipLaunchKernelGGL(restart_cells_kernel, dim3(m), dim3(64), 0, b->stream, A);
  HIPCHK(b, hipGetLastError());
  // the staging vectors are reused by the next call
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int cep_park_cells(cep_batch *b, const int *idx, int m)
{
  if (!b) return CEP_E_ARG;
  if (b->model != CEP_MODEL_ORD_CIPA && b->model != CEP_MODEL_TT06)
    return fail(b, CEP_E_ARG, "cells are parked one by one in ORd-CiPA and TT06 batches");
  if (b->pending) return fail(b, CEP_E_STATE, "a run is in flight: cep_sync or cep_query first");
  if (int rc = stage_cells(b, idx, m)) return rc;
  hipLaunchKernelGGL(park_cells_kernel, dim3((m + 63) / 64), dim3(64), 0, b->stream, m, b->d_stage_idx,
                     b->d_status);
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"

#include "autotune.hpp"

#include "mesh.hpp"
#include "tissue.hpp"

#include "io.hpp"

// ---- one fixed-dt step of every cell: computeRates(t) then solveAnalytical(dt), the step of the
// reference's fixed-dt loops (population_tn2006.cpp:145-155 for TT06, drug_sim.cpp:330-336 for
// ORd 2011).  ORd-CiPA's solveAnalytical is empty in the reference (:1113-1115); its step is the
// tissue reaction scheme (Rush-Larsen gates, backward-Euler Markov block, forward Euler rest).
template <typename R>
__global__ void __launch_bounds__(256) tt_step_kernel(int n, int pitch, double t, double dt, double *__restrict__ y,
                                                      const double *__restrict__ prm)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double s[tt::NS];
#pragma unroll
  for (int i = 0; i < tt::NS; ++i) s[i] = y[(size_t)i * pitch + c];
  const PrmAcc P{prm + c, pitch};
  double dvdt;
  tt::Currents cur;
  tt::rush_larsen<R>(t, dt, s, P, dvdt, &cur);
#pragma unroll
  for (int i = 0; i < tt::NS; ++i) y[(size_t)i * pitch + c] = s[i];
}

extern "C" {

int cep_step_rl(cep_batch *b, double t, double dt)
{
  if (!b || !(dt > 0) || !(t >= 0)) return fail(b, CEP_E_ARG, "step arguments");
  if (b->cap == 0) return fail(b, CEP_E_STATE, "cep_init_consts not called");
  const bool f64 = b->prec == CEP_PREC_FP64;
  if (b->model == CEP_MODEL_TT06) {
    const int blocks = (b->n + 255) / 256;
    if (f64) hipLaunchKernelGGL(tt_step_kernel<double>, dim3(blocks), dim3(256), 0, b->stream, b->n, b->pitch, t, dt, b->d_y, b->d_prm);
    else hipLaunchKernelGGL(tt_step_kernel<float>, dim3(blocks), dim3(256), 0, b->stream, b->n, b->pitch, t, dt, b->d_y, b->d_prm);
  } else if (b->model == CEP_MODEL_ORD2011) {
    const int blocks = (b->n + 63) / 64;
    if (f64) hipLaunchKernelGGL(cep::tis::tissue_react11<double>, dim3(blocks), dim3(64), 0, b->stream, b->n, b->pitch, t, dt, -1, b->d_y, b->d_prm,
                                (const int *)nullptr, b->pitch);
    else hipLaunchKernelGGL(cep::tis::tissue_react11<float>, dim3(blocks), dim3(64), 0, b->stream, b->n, b->pitch, t, dt, -1, b->d_y, b->d_prm,
                                (const int *)nullptr, b->pitch);
  } else {
    const int blocks = (b->n + 63) / 64;
    if (f64) hipLaunchKernelGGL(cep::tis::tissue_react<double>, dim3(blocks), dim3(64), 0, b->stream, b->n, b->pitch, t, dt, b->d_y, b->d_prm,
                                (const int *)nullptr, b->pitch);
    else hipLaunchKernelGGL(cep::tis::tissue_react<float>, dim3(blocks), dim3(64), 0, b->stream, b->n, b->pitch, t, dt, b->d_y, b->d_prm,
                            (const int *)nullptr, b->pitch);
  }
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"

// ---- computeRates through a fixed device array (no allocation on the call path)
//
// The staging of cep_compute_rates_static -- the states, rates and ALGEBRAIC rows of up to
// kStaticRatesCells cells of the widest model (ORd-CiPA: 52 + 52 + 200 doubles a cell), 4.98 MB --
// lives in one file-scope device array.  Each GPU the library runs on holds its own copy until the
// process ends.  Host code uses the array only through the device address hipGetSymbolAddress
// gives for the current device (the array's name is a host-side stand-in there), looked up on
// the first call on each device.  Every call on a GPU shares that GPU's array: calls that overlap
// on one GPU (from several host threads or streams) race.
namespace staticrates {
constexpr int kStaticRatesCells = 2048;
constexpr size_t kRowDoubles = (size_t)ord::NS + ord::NS + ORD_NALG;   // the widest model's row
constexpr size_t kScratchDoubles = (size_t)kStaticRatesCells * kRowDoubles;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex mu;
static double *dev_ptr[kMaxDevices];

static int scratch_for_current_device(cep_batch *b, double **out)
{
  int dev = 0;
  HIPCHK(b, hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(b, CEP_E_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(mu);
  if (!dev_ptr[dev]) {
    double *p = nullptr;
    HIPCHK(b, hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(b, CEP_E_HIP, "hipGetSymbolAddress gave no address for the scratch array");
    dev_ptr[dev] = p;
  }
  *out = dev_ptr[dev];
  return 0;
}

__global__ void __launch_bounds__(256) fill_zero(double *__restrict__ p, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
}  // namespace staticrates

extern "C" {

int cep_compute_rates_static(cep_batch *b, double t, const double *states, double *rates, double *algebraic)
{
  using namespace staticrates;
  if (!b || !states || !rates) return fail(b, CEP_E_ARG, "null array");
  if (b->n > kStaticRatesCells)
    return fail(b, CEP_E_ARG, "cep_compute_rates_static takes at most 2048 cells (its fixed device array); "
                              "use cep_compute_rates for larger batches");
  double *base = nullptr;
  if (int rc = scratch_for_current_device(b, &base)) return rc;
  const size_t nsz = (size_t)b->n * b->ns, asz = (size_t)b->n * b->nalg;
  double *dy = base, *df = base + nsz, *da = algebraic ? base + 2 * nsz : nullptr;
  if (da) {   // the kernels write only the ALGEBRAIC slots they compute: the rest must read zero
    hipLaunchKernelGGL(fill_zero, dim3((unsigned)((asz + 255) / 256)), dim3(256), 0, b->stream, da, asz);
    HIPCHK(b, hipGetLastError());
  }
  HIPCHK(b, hipMemcpyAsync(dy, states, nsz * sizeof(double), hipMemcpyHostToDevice, b->stream));
  constexpr int kRatesBlock = 64;
  const int blocks = (b->n + kRatesBlock - 1) / kRatesBlock;
  const bool f64 = b->prec == CEP_PREC_FP64;
  if (b->model == CEP_MODEL_ORD_CIPA) {
    if (f64) hipLaunchKernelGGL(ord_rates_kernel<double>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
    else hipLaunchKernelGGL(ord_rates_kernel<float>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
  } else if (b->model == CEP_MODEL_ORD2011) {
    if (f64) hipLaunchKernelGGL(ord11_rates_kernel<double>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
    else hipLaunchKernelGGL(ord11_rates_kernel<float>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
  } else {
    if (f64) hipLaunchKernelGGL(tt_rates_kernel<double>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
    else hipLaunchKernelGGL(tt_rates_kernel<float>, dim3(blocks), dim3(kRatesBlock), 0, b->stream, b->n, b->pitch, t, dy, df, da, b->d_prm);
  }
  HIPCHK(b, hipGetLastError());
  HIPCHK(b, hipMemcpyAsync(rates, df, nsz * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  if (da) HIPCHK(b, hipMemcpyAsync(algebraic, da, asz * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"
