// tissue.hpp -- monodomain tissue step on gfx950: ORd-CiPA reaction per node + FEM diffusion
// (a 2-D Q1 grid, or a mesh of quadrilaterals, hexahedra or tetrahedra in CSR) solved by
// Jacobi-preconditioned CG.  Included once, at the end of cep_gpu.hip
// (it uses the batch internals defined there).
//
// Reference: Cheart::cellular_reaction (heart.cpp:274-430) and Cfem::set_amat_bmat /
// assemble_Amat_Bmat / solve_Bmat (fem.cpp:628-661,721-763,705-718), Q1 shape functions and
// 2x2 Gauss points (fem.cpp:256-271,398-426).  Layout and kernels:
//   * nodes (nx+1) x (ny+1), index iy*(nx+1)+ix; the states of the node cells are the SoA
//     state planes of an ORd batch (plane 0 = Vm, read and written in place by the solver);
//   * A = M/dt + K and B = M/dt assembled on the host as nine FP64 coefficient planes
//     (stencil slot s = (dy+1)*3 + (dx+1)), zero outside the grid, rows padded to a multiple
//     of the tile width so that a tile row is whole 128 B lines;
//   * one CG iteration = two kernels on the grid: the SpMV q = A p on 32x32 node tiles with
//     the p tile and its halo staged in LDS, which also forms the direction p = z + b p_prev
//     while staging and runs the previous iteration's stop test, fused with the block partial
//     of p.q; then the vector update x += a p, r -= a q, z = D^-1 r fused with the partials of
//     r.z and z.z.  (A mesh in CSR keeps a separate direction kernel.)  The scalars a, b and
//     the stop test are re-reduced from the per-workgroup partials by every workgroup of the
//     next kernel in one fixed order (results do not depend on scheduling), so iterations run
//     without host round trips; the host checks the stop flag every few iterations.
// Algorithmic HBM bytes per CG iteration and node: SpMV 9*8 (A) + 8 (z) + 8 (p_prev) read,
// 8 (p) + 8 (q) written = 104; update 8*5 read (x, p, r, q, dinv) + 8*3 written = 64; total
// 168 B/node/iteration.
#pragma once

namespace cep {
namespace tis {

// SpMV node tile per workgroup (16 waves).  The next kernels re-reduce the per-workgroup
// partials of the previous one (every workgroup, the same fixed order), a cost that grows with
// the square of the workgroup count: 32 x 32 tiles (257 workgroups at 513^2 nodes) against
// 16 x 16 (1028) took a CG update from 14.4 to 12.6 us and a Strang step from 0.422 to 0.402 ms.
constexpr int TX = 32, TY = 32, TB = TX * TY;

struct Scal {            // device-resident CG state
  double rtol2;
  int done, iters, maxit;
};

// partial sums of the block reductions: [PQ | RZ(even) | RZ(odd) | ZZ | ZZ0] x MAXBLK
enum { P_PQ = 0, P_RZ0 = 1, P_RZ1 = 2, P_ZZ = 3, P_ZZ0 = 4, P_N = 5 };

// workgroup sum of NV values; the result is valid in thread 0
template <int NV>
CEP_D void block_sum(double (&v)[NV], double *sm)
{
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[wid * NV + k] = v[k];
  __syncthreads();
  // the TB/64 wave sums by a fixed shuffle tree in the first wave (not a serial chain of LDS
  // reads on one lane)
  static_assert(TB / 64 == 16, "sixteen wave sums");
  if (wid == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = lane < 16 ? sm[lane * NV + k] : 0.0;
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      v[k] = s;
    }
}

// the totals of NV partial arrays (part[k], cnt[k] values each, written by an earlier kernel),
// computed identically by every workgroup -- fixed strided order per thread, fixed tree -- in
// one reduction, and broadcast through LDS
template <int NV>
CEP_D void grid_totals(const double *const (&part)[NV], const int (&cnt)[NV], double (&tot)[NV], double *sm,
                       double *bc)
{
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    v[k] = 0.0;
    for (int i = threadIdx.x; i < cnt[k]; i += TB) v[k] += part[k][i];
  }
  block_sum<NV>(v, sm);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) bc[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = bc[k];
}

struct Grid {
  int nxn, nyn, nn;
  int nblk;      // SpMV workgroups: 2-D tiles of the grid, or TB nodes of a mesh
  int nblk1;     // vector-kernel workgroups: TB consecutive nodes
  int cp;        // grid: row pitch of the coefficient planes (nxn rounded up to TX); plane = cp * nyn
  int lin;       // 1: a mesh (CSR SpMV)
  int maxblk;    // stride of the partial arrays
};

CEP_D int block_id() { return blockIdx.y * gridDim.x + blockIdx.x; }

// the node of this thread in the vector kernels: TB consecutive nodes per workgroup
CEP_D int node_1d(const Grid &g)
{
  const int i = blockIdx.x * TB + threadIdx.x;
  return i < g.nn ? i : -1;
}

// The grid SpMV q = A p over a 32 x 32 tile, the tile of p and its one-node halo staged in LDS
// (zero outside the grid).  Coefficient plane s at A[s * cp * nyn + iy * cp + ix]: with the
// row pitch a multiple of the tile width, every tile row of a plane is two whole 128 B lines.
// DIR: the CG direction of iteration k >= 1 is formed while staging, p_k = z + b p_(k-1) with
// b = RZ(k) / RZ(k-1), read from pin and written to pout (two buffers: no workgroup overwrites
// the previous direction that another one stages as its halo); the stop test of iteration k-1
// comes first (||z|| <= rtol ||z0||, PETSc KSPCG's preconditioned norm, or k == maxit).  Every
// workgroup computes the same totals in the same order, so every halo value equals what its
// owner writes, and all workgroups take the same decision.  The partial of p.q goes to PQ.
// After the stop flag (launches queued past the end of the solve return at once), the loads of
// the tile and the coefficients are issued before the reduction of the partials (a scheduling
// barrier keeps them there), so their memory latencies overlap instead of following one another.
constexpr int NSTAGE = ((TX + 2) * (TY + 2) + TB - 1) / TB;   // tile elements per thread
template <bool DIR>
__global__ void __launch_bounds__(TB) cg_spmv(Grid g, const double *__restrict__ A, const double *__restrict__ z,
                                              const double *__restrict__ pin, double *__restrict__ pout,
                                              double *__restrict__ q, double *__restrict__ part, Scal *S, int k)
{
  __shared__ double tile[TY + 2][TX + 2];
  __shared__ double sm[TB / 64 * 4];
  __shared__ double bc[4];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int ix = blockIdx.x * TX + tx, iy = blockIdx.y * TY + ty;
  const bool mine = ix < g.nxn && iy < g.nyn;
  const bool form = DIR && k > 0;
  if (DIR && S->done) return;                    // launches queued past the end of the solve
  double pv[NSTAGE], zv[NSTAGE], a[9];
#pragma unroll
  for (int u = 0; u < NSTAGE; ++u) {
    const int m = threadIdx.x + u * TB;
    const int hx = m % (TX + 2), hy = m / (TX + 2);
    const int gx = blockIdx.x * TX + hx - 1, gy = blockIdx.y * TY + hy - 1;
    const bool in = m < (TX + 2) * (TY + 2) && gx >= 0 && gx < g.nxn && gy >= 0 && gy < g.nyn;
    const int n = in ? gy * g.nxn + gx : 0;
    pv[u] = in ? pin[n] : 0.0;
    zv[u] = (in && form) ? z[n] : 0.0;
  }
  const size_t plane = (size_t)g.cp * g.nyn, o = mine ? (size_t)iy * g.cp + ix : 0;
#pragma unroll
  for (int sl = 0; sl < 9; ++sl) a[sl] = mine ? A[sl * plane + o] : 0.0;
  __builtin_amdgcn_sched_barrier(0);
  double bt = 0.0;
  if (form) {
    const double *pa[4] = { part + P_ZZ * g.maxblk, part + P_ZZ0 * g.maxblk,
                            part + ((k & 1) ? P_RZ1 : P_RZ0) * g.maxblk, part + ((k & 1) ? P_RZ0 : P_RZ1) * g.maxblk };
    const int ca[4] = { g.nblk1, g.nblk1, g.nblk1, g.nblk1 };
    double tot[4];
    grid_totals<4>(pa, ca, tot, sm, bc);     // zz, zz0, RZ(k), RZ(k-1)
    if (tot[0] <= S->rtol2 * tot[1] || k >= S->maxit) {
      if (block_id() == 0 && threadIdx.x == 0) { S->done = 1; S->iters = (tot[1] > 0.0) ? k : 0; }
      return;
    }
    bt = tot[2] / tot[3];
    if (block_id() == 0 && threadIdx.x == 0) S->iters = k;
  }
#pragma unroll
  for (int u = 0; u < NSTAGE; ++u) {
    const int m = threadIdx.x + u * TB;
    if (m < (TX + 2) * (TY + 2)) tile[m / (TX + 2)][m % (TX + 2)] = form ? zv[u] + bt * pv[u] : pv[u];
  }
  __syncthreads();
  double v[1] = { 0.0 };
  if (mine) {
    const int i = iy * g.nxn + ix;
    double s = 0.0;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) s += a[(dy + 1) * 3 + (dx + 1)] * tile[ty + 1 + dy][tx + 1 + dx];
    q[i] = s;
    const double pi = tile[ty + 1][tx + 1];
    if (form) pout[i] = pi;
    v[0] = pi * s;
  }
  if (!DIR) return;                              // plain product (b = B v)
  block_sum<1>(v, sm);
  if (threadIdx.x == 0) part[P_PQ * g.maxblk + block_id()] = v[0];
}

// q = A p for a mesh matrix in CSR (columns ascending in each row); partial p.q (mode 1)
__global__ void __launch_bounds__(TB) cg_spmv_csr(Grid g, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const double *__restrict__ val, const double *__restrict__ p,
                                                  double *__restrict__ q, double *__restrict__ part, const Scal *S,
                                                  int mode)
{
  if (mode && S->done) return;
  __shared__ double sm[TB / 64 * 2];
  const int i = node_1d(g);
  double v[1] = { 0.0 };
  if (i >= 0) {
    double s = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) s += val[j] * p[ci[j]];
    q[i] = s;
    v[0] = p[i] * s;
  }
  if (mode == 0) return;                        // plain product (b = B v)
  block_sum<1>(v, sm);
  if (threadIdx.x == 0) part[P_PQ * g.maxblk + block_id()] = v[0];
}

// CG start from x = 0: r = b, z = D^-1 r, p = z; partials r.z (into RZ(even)) and z.z (ZZ0);
// the solve's scalars reset here (the later kernels of the solve read them), not by a copy
__global__ void __launch_bounds__(TB) cg_init(Grid g, const double *__restrict__ b, const double *__restrict__ dinv,
                                              double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
                                              double *__restrict__ p, double *__restrict__ part, Scal *S, double rtol2,
                                              int maxit)
{
  __shared__ double sm[TB / 64 * 2];
  if (block_id() == 0 && threadIdx.x == 0) { S->rtol2 = rtol2; S->done = 0; S->iters = 0; S->maxit = maxit; }
  const int i = node_1d(g);
  double v[2] = { 0.0, 0.0 };
  if (i >= 0) {
    const double ri = b[i], zi = dinv[i] * ri;
    x[i] = 0.0; r[i] = ri; z[i] = zi; p[i] = zi;
    v[0] = ri * zi; v[1] = zi * zi;
  }
  block_sum<2>(v, sm);
  if (threadIdx.x == 0) {
    part[P_RZ0 * g.maxblk + block_id()] = v[0];
    part[P_ZZ0 * g.maxblk + block_id()] = v[1];
  }
}

// iteration k: a = RZ(k) / PQ; x += a p, r -= a q, z = D^-1 r; partials of r.z (RZ(k+1)) and
// z.z; TB consecutive nodes per workgroup (whole, aligned lines)
__global__ void __launch_bounds__(TB) cg_update(Grid g, const double *__restrict__ p, const double *__restrict__ q,
                                                const double *__restrict__ dinv, double *__restrict__ x,
                                                double *__restrict__ r, double *__restrict__ z,
                                                double *__restrict__ part, const Scal *S, int k)
{
  __shared__ double sm[TB / 64 * 2];
  __shared__ double bc[2];
  if (S->done) return;                           // launches queued past the end of the solve
  // the vector loads first, then the reduction (their latencies overlap)
  const int i = node_1d(g);
  const int j = i >= 0 ? i : 0;
  const double pj = p[j], qj = q[j], xj = x[j], rj = r[j], dj = dinv[j];
  __builtin_amdgcn_sched_barrier(0);
  const double *pa[2] = { part + ((k & 1) ? P_RZ1 : P_RZ0) * g.maxblk, part + P_PQ * g.maxblk };
  const int ca[2] = { g.nblk1, g.nblk };
  double tot[2];
  grid_totals<2>(pa, ca, tot, sm, bc);
  const double a = (tot[1] != 0.0) ? tot[0] / tot[1] : 0.0;    // b = 0: x stays 0
  double v[2] = { 0.0, 0.0 };
  if (i >= 0) {
    x[i] = xj + a * pj;
    const double ri = rj - a * qj;
    const double zi = dj * ri;
    r[i] = ri; z[i] = zi;
    v[0] = ri * zi; v[1] = zi * zi;
  }
  block_sum<2>(v, sm);
  if (threadIdx.x == 0) {
    part[((k & 1) ? P_RZ0 : P_RZ1) * g.maxblk + block_id()] = v[0];
    part[P_ZZ * g.maxblk + block_id()] = v[1];
  }
}

// mesh tissue, iteration k: stop test ||z|| <= rtol ||z0|| (preconditioned norm, as PETSc
// KSPCG) or k+1 == maxit; else b = RZ(k+1) / RZ(k), p = z + b p
__global__ void __launch_bounds__(TB) cg_dir(Grid g, const double *__restrict__ z, double *__restrict__ p,
                                             const double *__restrict__ part, Scal *S, int k)
{
  if (S->done) return;
  __shared__ double sm[TB / 64 * 4];
  __shared__ double bc[4];
  const double *pa[4] = { part + P_ZZ * g.maxblk, part + P_ZZ0 * g.maxblk,
                          part + ((k & 1) ? P_RZ0 : P_RZ1) * g.maxblk, part + ((k & 1) ? P_RZ1 : P_RZ0) * g.maxblk };
  const int ca[4] = { g.nblk1, g.nblk1, g.nblk1, g.nblk1 };
  double tot[4];
  grid_totals<4>(pa, ca, tot, sm, bc);        // zz, zz0, RZ(k+1), RZ(k)
  if (tot[0] <= S->rtol2 * tot[1] || k + 1 >= S->maxit) {
    if (block_id() == 0 && threadIdx.x == 0) { S->done = 1; S->iters = (tot[1] > 0.0) ? k + 1 : 0; }
    return;
  }
  const double bt = tot[2] / tot[3];
  const int i = node_1d(g);
  if (i >= 0) p[i] = z[i] + bt * p[i];
  if (block_id() == 0 && threadIdx.x == 0) S->iters = k + 1;
}

// S2 of the S1-S2 protocol (Cheart::apply_s2, heart.cpp:475-501): one workgroup; the target
// node's Vm arms the reset when above -45 mV and fires it when below again; the reset writes
// the S2 nodes' initial states (s2init: count x ns, node-major) and records the time.
struct S2State {
  int up, down, done, pad;
  double time;
};
__global__ void __launch_bounds__(256) tissue_s2(int target, int pitch, int ns, double t, double *__restrict__ y,
                                                 S2State *st, const int *__restrict__ nodes, int count,
                                                 const double *__restrict__ s2init)
{
  __shared__ int fire;
  if (threadIdx.x == 0) {
    fire = 0;
    if (!st->done) {
      const double v = y[target];
      if (v > -45.0) st->up = 1;
      if (st->up && v < -45.0) st->down = 1;
      if (st->down) { fire = 1; st->done = 1; st->time = t; }
    }
  }
  __syncthreads();
  if (!fire) return;
  for (int k = threadIdx.x; k < count * ns; k += blockDim.x) {
    const int n = k / ns, j = k % ns;
    y[(size_t)j * pitch + nodes[n]] = s2init[k];
  }
}

// reaction step of every node over h (the reference's solveAnalytical role; ORd-CiPA has
// none there, :1113-1115): Rush-Larsen for the Hodgkin-Huxley gates and the Jrel
// relaxations, backward Euler for the Markov IKr block (linear in its states; drug binding
// and trapping make it stiff: rates up to Kmax*Ku), forward Euler for V, concentrations,
// CaMK and the charge accumulators.  All rates at the start of the step.
// at most 256 registers, two waves per SIMD (measured, 263,169 nodes in FP64: 108 us per
// launch against 116 us with 326 registers and one wave)
// Node parameters come from a table of the distinct parameter rows (`ptab`, row r's slot k at
// ptab[k * nclass + r]) and each node's row index (`pcls`): the nodes of a tissue share a few
// rows (cell type, drug, paced or not), so a node reads one index instead of its 30 doubles,
// and the table stays in cache.  pcls null: every cell its own row (ptab = the batch's SoA
// parameters, nclass = its pitch).
template <typename R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) tissue_react(int n, int pitch, double t, double h, double *__restrict__ y,
                                                   const double *__restrict__ ptab, const int *__restrict__ pcls,
                                                   int nclass)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double s[ord::NS], f[ord::NS], a[ord::NS];
#pragma unroll
  for (int i = 0; i < ord::NS; ++i) { s[i] = y[(size_t)i * pitch + c]; a[i] = 0.0; }
  ord::Currents cur;
  const PrmAcc P{ptab + (pcls ? pcls[c] : c), nclass};
  ord::rhs<R>(t, s, f, P, &cur, a);
  // Markov block: (I - h A) S1 = S0, no pivoting (an M-matrix)
  double M[9][9], b[9];
  ord::markov_matrix<R>(s[0], P, M);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
#pragma unroll
    for (int j = 0; j < 9; ++j) M[i][j] = (i == j ? 1.0 : 0.0) - h * M[i][j];
    b[i] = s[34 + i];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double pinv = 1.0 / M[k][k];
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      const double l = M[i][k] * pinv;
#pragma unroll
      for (int j = k + 1; j < 9; ++j) M[i][j] -= l * M[k][j];
      b[i] -= l * b[k];
    }
  }
#pragma unroll
  for (int i = 8; i >= 0; --i) {
#pragma unroll
    for (int j = i + 1; j < 9; ++j) b[i] -= M[i][j] * b[j];
    b[i] /= M[i][i];
  }
#pragma unroll
  for (int i = 0; i < ord::NS; ++i) {
    double v;
    if (i >= 34 && i <= 42) {
      v = b[i - 34];
    } else if (a[i] > 0.0) {
      // Rush-Larsen: s + f (1 - exp(-h a)) / a = s + f h phi(h a), phi(x) = (1 - e^-x) / x
      if constexpr (sizeof(R) == sizeof(float)) {
        // FP32 rates: phi at FP32 relative accuracy (its Taylor series to x^4 below x = 1/32,
        // truncation x^5/720 < 1e-10), so the FP64 state takes the step's change without the
        // rounding of exp(-x) to FP32 near 1
        const float x = (float)(h * a[i]);
        const float ph = x < 0.03125f ? 1.f - x * (0.5f - x * ((1.f / 6.f) - x * ((1.f / 24.f) - x * (1.f / 120.f))))
                                      : (1.f - __expf(-x)) * __builtin_amdgcn_rcpf(x);
        v = s[i] + f[i] * h * (double)ph;
      } else {
        v = s[i] + f[i] * (-expm1(-h * a[i]) / a[i]);
      }
    } else {
      v = s[i] + h * f[i];
    }
    y[(size_t)i * pitch + c] = v;
  }
}

// reaction step of every ORd 2011 node, the reference tissue driver's own scheme
// (Cheart::cellular_reaction, heart.cpp:341-368): computeRates with the TISSUE stimulus
// (isS1 ? amp : 0; unpaced nodes carry amp = 0) then solveAnalytical(h).
template <typename R>
__global__ void __launch_bounds__(64) tissue_react11(int n, int pitch, double t, double h, int stim_on,
                                                     double *__restrict__ y, const double *__restrict__ ptab,
                                                     const int *__restrict__ pcls, int nclass)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double s[ord11::NS], f[ord11::NS];
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) s[i] = y[(size_t)i * pitch + c];
  const PrmAcc P{ptab + (pcls ? pcls[c] : c), nclass};
  ord11::Gates g;
  ord11::Currents cur;
  ord11::rhs<R>(t, s, f, P, cur, g, stim_on);
  ord11::rl_step<R>(h, s, f, g);
#pragma unroll
  for (int i = 0; i < ord11::NS; ++i) y[(size_t)i * pitch + c] = s[i];
}

}  // namespace tis
}  // namespace cep

struct cep_tissue {
  int model = CEP_MODEL_ORD_CIPA;
  cep_batch *cells = nullptr;
  int nx = 0, ny = 0, nxn = 0, nyn = 0, nn = 0;
  double hx = 0, hy = 0, t = 0, dt_built = -1;
  std::string err;
  std::vector<double> sxx, sxy, syy;       // per element
  // mesh tissue: element kind (quadrilateral, hexahedron, tetrahedron), node coordinates,
  // connectivity, tensor per element and the CSR pattern of A and B
  bool mesh = false;
  int ekind = 0, npe = 4, dim = 2;
  int ne = 0, nnz = 0;
  std::vector<double> xyz;                 // nn x 3
  std::vector<int> conn;                   // ne x npe
  std::vector<double> dten;                // ne x 6: xx, xy, xz, yy, yz, zz
  std::vector<int> rp, ci;                 // CSR rows / columns (ascending)
  int *d_rp = nullptr, *d_ci = nullptr;
  std::vector<int> stim_nodes;
  double stim_scale = 1.0, stim_bcl = 1000.0, stim_start = 0.0, stim_dur = 0.5;
  // S1-S2 protocol
  int npace1 = 0, s2_count = 0, s2_target = 0;
  bool s1_stopped = false;
  int *d_s2nodes = nullptr;
  double *d_s2init = nullptr;
  cep::tis::S2State *d_s2 = nullptr;
  double *d_A = nullptr, *d_B = nullptr, *d_dinv = nullptr;
  double *d_b = nullptr, *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_q = nullptr, *d_part = nullptr;
  double *d_p2 = nullptr;                  // grid: the second direction buffer (directions alternate)
  // node parameters as a table of distinct rows and a row index per node
  double *d_ptab = nullptr;
  int *d_pcls = nullptr;
  int nclass = 0, ptab_cap = 0;
  int cp = 0;                              // grid: row pitch of the coefficient planes
  cep::tis::Scal *d_S = nullptr;
  std::vector<void *> allocs;
  std::vector<hipEvent_t> ev;              // pairs for kernel timing
  double tsum[3] = { 0, 0, 0 };
  long long tcount[3] = { 0, 0, 0 };
};

static int tfail(cep_tissue *t, int code, const std::string &msg)
{
  if (t) t->err = msg;
  return code;
}

#define TCHK(t, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return tfail(t, CEP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int talloc(cep_tissue *t, T **p, size_t count)
{
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T) + 64);
  if (e == hipSuccess) e = hipMemsetAsync(q, 0, count * sizeof(T) + 64, t->cells->stream);
  if (e != hipSuccess) return tfail(t, CEP_E_HIP, std::string("tissue alloc: ") + hipGetErrorString(e));
  t->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

// global 9-point stencils of A = M/dt + K and B = M/dt from the Q1 element matrices
static void tissue_assemble(cep_tissue *t, double dt, std::vector<double> &A, std::vector<double> &Bm,
                            std::vector<double> &dinv)
{
  const int nn = t->nn, nxn = t->nxn;
  A.assign((size_t)9 * nn, 0.0);
  Bm.assign((size_t)9 * nn, 0.0);
  // element node order (0,0), (1,0), (1,1), (0,1) as the reference's set_Phi_dPhiS_square
  const int ox[4] = { 0, 1, 1, 0 }, oy[4] = { 0, 0, 1, 1 };
  const double g[2] = { -0.57735, 0.57735 };             // the reference's Gauss points (fem.cpp:256-270)
  const double gs[4] = { g[0], g[1], g[1], g[0] }, gt[4] = { g[0], g[0], g[1], g[1] };
  const double hx = t->hx, hy = t->hy, detJ = 0.25 * hx * hy;
  double Me[16] = { 0 }, Ks[3][16] = { { 0 } };            // mass and the three tensor parts of K
  for (int gp = 0; gp < 4; ++gp) {
    const double s = gs[gp], u = gt[gp];
    const double phi[4] = { (1. - s) * (1. - u) / 4., (1. + s) * (1. - u) / 4., (1. + s) * (1. + u) / 4., (1. - s) * (1. + u) / 4. };
    const double dps[4] = { -(1. - u) / 4., (1. - u) / 4., (1. + u) / 4., -(1. + u) / 4. };
    const double dpt[4] = { -(1. - s) / 4., -(1. + s) / 4., (1. + s) / 4., (1. - s) / 4. };
    double ddx[4], ddy[4];
    for (int a = 0; a < 4; ++a) { ddx[a] = dps[a] * (2.0 / hx); ddy[a] = dpt[a] * (2.0 / hy); }
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        Me[j * 4 + k] += phi[j] * phi[k] * detJ;
        Ks[0][j * 4 + k] += ddx[j] * ddx[k] * detJ;                          // sxx part
        Ks[1][j * 4 + k] += (ddx[j] * ddy[k] + ddy[j] * ddx[k]) * detJ;      // sxy part
        Ks[2][j * 4 + k] += ddy[j] * ddy[k] * detJ;                          // syy part
      }
  }
  for (int ey = 0; ey < t->ny; ++ey)
    for (int ex = 0; ex < t->nx; ++ex) {
      const int e = ey * t->nx + ex;
      for (int j = 0; j < 4; ++j) {
        const int i = (ey + oy[j]) * nxn + (ex + ox[j]);
        for (int k = 0; k < 4; ++k) {
          const int dx = ox[k] - ox[j], dy = oy[k] - oy[j];
          const int slot = (dy + 1) * 3 + (dx + 1);
          const double kk = t->sxx[e] * Ks[0][j * 4 + k] + t->sxy[e] * Ks[1][j * 4 + k] + t->syy[e] * Ks[2][j * 4 + k];
          A[(size_t)slot * nn + i] += Me[j * 4 + k] / dt + kk;
          Bm[(size_t)slot * nn + i] += Me[j * 4 + k] / dt;
        }
      }
    }
  dinv.resize(nn);
  for (int i = 0; i < nn; ++i) dinv[i] = 1.0 / A[(size_t)4 * nn + i];
}

// CSR values of A = M/dt + K and B = M/dt on an S4 mesh.  Each element's matrices follow
// Cfem::set_amat_bmat (fem.cpp:628-661): at the reference's four Gauss points (+-0.57735) the
// Jacobian of the bilinear map of the element's nodes, its determinant and the physical
// gradients of the shape functions; K uses the element's conductivity tensor.
// Element tables of the mesh tissue (fem.cpp: set_gpt_* :209-270, set_Phi_dPhiS_* :398-479):
// shape functions and their derivatives in reference coordinates at the Gauss points, with
// the reference's points (+-0.57735 for the quadrilateral and the hexahedron, the 4-point
// rule (5 -+ sqrt 5)/20 for the tetrahedron) and node orders.
enum ElemKind { EK_QUAD4 = 0, EK_HEX8 = 1, EK_TET4 = 2 };

struct ElemTable {
  int npe = 0, dim = 0, ng = 0;
  double w[8];                 // quadrature weight per point
  double phi[8][8];            // [point][node]
  double dphi[8][8][3];        // [point][node][reference coordinate]
};

static ElemTable elem_table(int kind)
{
  ElemTable E{};
  const double a = 0.57735;
  if (kind == EK_QUAD4) {
    E.npe = 4; E.dim = 2; E.ng = 4;
    const double gs[4] = { -a, a, a, -a }, gt[4] = { -a, -a, a, a };
    const int sx[4] = { -1, 1, 1, -1 }, sy[4] = { -1, -1, 1, 1 };
    for (int g = 0; g < 4; ++g) {
      E.w[g] = 1.0;
      for (int n = 0; n < 4; ++n) {
        E.phi[g][n] = (1 + sx[n] * gs[g]) * (1 + sy[n] * gt[g]) / 4.;
        E.dphi[g][n][0] = sx[n] * (1 + sy[n] * gt[g]) / 4.;
        E.dphi[g][n][1] = sy[n] * (1 + sx[n] * gs[g]) / 4.;
      }
    }
  } else if (kind == EK_HEX8) {
    // points in the order of the reference's ix, iy, iz loops; nodes: the bottom face
    // counter-clockwise, then the top face
    E.npe = 8; E.dim = 3; E.ng = 8;
    const int sx[8] = { -1, 1, 1, -1, -1, 1, 1, -1 }, sy[8] = { -1, -1, 1, 1, -1, -1, 1, 1 };
    const int sz[8] = { -1, -1, -1, -1, 1, 1, 1, 1 };
    int g = 0;
    for (int ix = 0; ix < 2; ++ix)
      for (int iy = 0; iy < 2; ++iy)
        for (int iz = 0; iz < 2; ++iz, ++g) {
          const double p[3] = { ix ? a : -a, iy ? a : -a, iz ? a : -a };
          E.w[g] = 1.0;
          for (int n = 0; n < 8; ++n) {
            const double fx = 1 + sx[n] * p[0], fy = 1 + sy[n] * p[1], fz = 1 + sz[n] * p[2];
            E.phi[g][n] = fx * fy * fz / 8.;
            E.dphi[g][n][0] = sx[n] * fy * fz / 8.;
            E.dphi[g][n][1] = sy[n] * fx * fz / 8.;
            E.dphi[g][n][2] = sz[n] * fx * fy / 8.;
          }
        }
  } else {
    // linear tetrahedron: phi = (1-s-t-z, s, t, z).  The reference sums its four points with
    // weight 1, which scales both A and B by 24 and leaves A x = B v unchanged; the weight
    // 1/24 here makes the mass matrix integrate to the volume.
    E.npe = 4; E.dim = 3; E.ng = 4;
    const double lo = (5. - sqrt(5.)) / 20., hi = (5. + 3. * sqrt(5.)) / 20.;
    const double P[4][3] = { { lo, lo, lo }, { lo, lo, hi }, { lo, hi, lo }, { hi, lo, lo } };
    for (int g = 0; g < 4; ++g) {
      E.w[g] = 1.0 / 24.0;
      E.phi[g][0] = 1. - P[g][0] - P[g][1] - P[g][2];
      for (int n = 1; n < 4; ++n) E.phi[g][n] = P[g][n - 1];
      for (int c = 0; c < 3; ++c) {
        E.dphi[g][0][c] = -1.;
        for (int n = 1; n < 4; ++n) E.dphi[g][n][c] = (n - 1 == c) ? 1. : 0.;
      }
    }
  }
  return E;
}

// Element mass and conduction matrices (npe x npe) of one element at coordinates X (npe x 3)
// with the tensor D = (xx, xy, xz, yy, yz, zz), as Cfem::set_amat_bmat (fem.cpp:628-661):
// per Gauss point J = dPhiS^T X, grad phi = J^-1 dphi/d(reference), then
// Me += w phi_j phi_k det J and Ke += w grad phi_j . D grad phi_k det J.  Returns the smallest
// det J over the points (a folded or inverted element has one <= 0).
static double elem_matrices(const ElemTable &E, const double (*X)[3], const double *D, double *Me, double *Ke)
{
  const int n = E.npe, d = E.dim;
  const double Dm[3][3] = { { D[0], D[1], D[2] }, { D[1], D[3], D[4] }, { D[2], D[4], D[5] } };
  for (int q = 0; q < n * n; ++q) { Me[q] = 0.0; Ke[q] = 0.0; }
  double detmin = 1e300;
  for (int g = 0; g < E.ng; ++g) {
    double J[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };   // rows: reference coord., columns: x, y, z
    for (int r = 0; r < d; ++r)
      for (int c = 0; c < d; ++c)
        for (int a = 0; a < n; ++a) J[r][c] += E.dphi[g][a][r] * X[a][c];
    double inv[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } }, det;
    if (d == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      inv[0][0] = J[1][1] / det; inv[0][1] = -J[0][1] / det;
      inv[1][0] = -J[1][0] / det; inv[1][1] = J[0][0] / det;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                   c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      // inverse = adjugate / det (the adjugate is the transposed cofactor matrix)
      inv[0][0] = c00 / det;
      inv[1][0] = c01 / det;
      inv[2][0] = c02 / det;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
    detmin = det < detmin ? det : detmin;
    double gr[8][3] = {};
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < d; ++c)
        for (int r = 0; r < d; ++r) gr[a][c] += inv[c][r] * E.dphi[g][a][r];
    const double wd = E.w[g] * det;
    for (int j = 0; j < n; ++j) {
      double Dg[3] = { 0, 0, 0 };
      for (int c = 0; c < d; ++c)
        for (int r = 0; r < d; ++r) Dg[c] += Dm[c][r] * gr[j][r];
      for (int k = 0; k < n; ++k) {
        Me[j * n + k] += E.phi[g][j] * E.phi[g][k] * wd;
        double kk = 0.0;
        for (int c = 0; c < d; ++c) kk += Dg[c] * gr[k][c];
        Ke[j * n + k] += kk * wd;
      }
    }
  }
  return detmin;
}

static void tissue_assemble_mesh(cep_tissue *t, double dt, std::vector<double> &A, std::vector<double> &Bm,
                                 std::vector<double> &dinv)
{
  A.assign(t->nnz, 0.0);
  Bm.assign(t->nnz, 0.0);
  const ElemTable E = elem_table(t->ekind);
  const int n = E.npe;
  auto slot = [&](int i, int j) {             // CSR position of (i, j)
    int lo = t->rp[i], hi = t->rp[i + 1] - 1;
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (t->ci[mid] < j) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  double Me[64], Ke[64], X[8][3];
  for (int e = 0; e < t->ne; ++e) {
    const int *nd = &t->conn[(size_t)e * n];
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < 3; ++c) X[a][c] = t->xyz[(size_t)nd[a] * 3 + c];
    elem_matrices(E, X, &t->dten[(size_t)e * 6], Me, Ke);
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        const int q = slot(nd[j], nd[k]);
        A[q] += Me[j * n + k] / dt + Ke[j * n + k];
        Bm[q] += Me[j * n + k] / dt;
      }
  }
  dinv.resize(t->nn);
  for (int i = 0; i < t->nn; ++i) dinv[i] = 1.0 / A[slot(i, i)];
}

static int tissue_build(cep_tissue *t, double dt)
{
  if (t->dt_built == dt) return 0;
  std::vector<double> A, Bm, dinv;
  if (t->mesh) tissue_assemble_mesh(t, dt, A, Bm, dinv);
  else tissue_assemble(t, dt, A, Bm, dinv);
  hipStream_t s = t->cells->stream;
  if (!t->mesh) {
    // nine planes with the row pitch padded to the tile width (tile rows on whole lines)
    const size_t plane = (size_t)t->cp * t->nyn;
    std::vector<double> Ap(9 * plane, 0.0), Bp(9 * plane, 0.0);
    for (int sl = 0; sl < 9; ++sl)
      for (int iy = 0; iy < t->nyn; ++iy)
        for (int ix = 0; ix < t->nxn; ++ix) {
          const size_t from = (size_t)sl * t->nn + (size_t)iy * t->nxn + ix, to = sl * plane + (size_t)iy * t->cp + ix;
          Ap[to] = A[from];
          Bp[to] = Bm[from];
        }
    A.swap(Ap);
    Bm.swap(Bp);
  }
  TCHK(t, hipMemcpyAsync(t->d_A, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice, s));
  TCHK(t, hipMemcpyAsync(t->d_B, Bm.data(), Bm.size() * sizeof(double), hipMemcpyHostToDevice, s));
  TCHK(t, hipMemcpyAsync(t->d_dinv, dinv.data(), dinv.size() * sizeof(double), hipMemcpyHostToDevice, s));
  TCHK(t, hipStreamSynchronize(s));
  t->dt_built = dt;
  return 0;
}

static void tissue_timed_begin(cep_tissue *t, int slot)
{
  (void)hipEventRecord(t->ev[2 * slot], t->cells->stream);
}
static void tissue_timed_end(cep_tissue *t, int slot)
{
  (void)hipEventRecord(t->ev[2 * slot + 1], t->cells->stream);
}
static void tissue_timed_collect(cep_tissue *t, int slot)
{
  float ms = 0.f;
  if (hipEventSynchronize(t->ev[2 * slot + 1]) == hipSuccess &&
      hipEventElapsedTime(&ms, t->ev[2 * slot], t->ev[2 * slot + 1]) == hipSuccess) {
    t->tsum[slot] += ms;
    t->tcount[slot] += 1;
  }
}

static cep::tis::Grid tissue_grid(const cep_tissue *t)
{
  using namespace cep::tis;
  Grid g{};
  g.nn = t->nn;
  g.nblk1 = (t->nn + TB - 1) / TB;
  if (t->mesh) {
    g.lin = 1;
    g.nblk = g.nblk1;
  } else {
    g.nxn = t->nxn; g.nyn = t->nyn; g.cp = t->cp;
    g.nblk = ((t->nxn + TX - 1) / TX) * ((t->nyn + TY - 1) / TY);
  }
  g.maxblk = std::max(g.nblk, g.nblk1);
  return g;
}

// one diffusion solve A x = B v with v = x on entry (x: device vector of nn).  Grid: two
// kernels per iteration (the SpMV with the direction update and the previous stop test, then
// the vector update); mesh: three (CSR SpMV, vector update, direction).  The host reads the
// stop flag every CHUNK iterations.
static int tissue_solve(cep_tissue *t, double dt, double *x, double rtol, int maxit, int *iters)
{
  using namespace cep::tis;
  if (int rc = tissue_build(t, dt)) return rc;
  hipStream_t s = t->cells->stream;
  const Grid g = tissue_grid(t);
  const dim3 g2 = t->mesh ? dim3(g.nblk) : dim3((t->nxn + TX - 1) / TX, (t->nyn + TY - 1) / TY);
  const dim3 g1(g.nblk1), blk(TB);
  if (t->mesh)
    hipLaunchKernelGGL(cg_spmv_csr, g1, blk, 0, s, g, t->d_rp, t->d_ci, t->d_B, x, t->d_b, t->d_part, t->d_S, 0);
  else
    hipLaunchKernelGGL(cg_spmv<false>, g2, blk, 0, s, g, t->d_B, nullptr, x, nullptr, t->d_b, t->d_part, t->d_S, 0);
  hipLaunchKernelGGL(cg_init, g1, blk, 0, s, g, t->d_b, t->d_dinv, x, t->d_r, t->d_z, t->d_p, t->d_part, t->d_S,
                     rtol * rtol, maxit);
  TCHK(t, hipGetLastError());
  constexpr int CHUNK = 4;
  Scal h{};
  double *pb[2] = { t->d_p, t->d_p2 };
  if (t->mesh) {
    for (int it = 0; it < maxit; it += CHUNK) {
      for (int k = it; k < it + CHUNK && k < maxit; ++k) {
        // time the first iteration of each chunk (later ones may exit early once converged)
        const bool tm = k == it;
        if (tm) tissue_timed_begin(t, 1);
        hipLaunchKernelGGL(cg_spmv_csr, g1, blk, 0, s, g, t->d_rp, t->d_ci, t->d_A, t->d_p, t->d_q, t->d_part,
                           t->d_S, 1);
        if (tm) { tissue_timed_end(t, 1); tissue_timed_begin(t, 2); }
        hipLaunchKernelGGL(cg_update, g1, blk, 0, s, g, t->d_p, t->d_q, t->d_dinv, x, t->d_r, t->d_z, t->d_part,
                           t->d_S, k);
        if (tm) tissue_timed_end(t, 2);
        hipLaunchKernelGGL(cg_dir, g1, blk, 0, s, g, t->d_z, t->d_p, t->d_part, t->d_S, k);
      }
      TCHK(t, hipGetLastError());
      TCHK(t, hipMemcpyAsync(&h, t->d_S, sizeof(Scal), hipMemcpyDeviceToHost, s));
      TCHK(t, hipStreamSynchronize(s));
      tissue_timed_collect(t, 1);
      tissue_timed_collect(t, 2);
      if (h.done) break;
    }
  } else {
    // SpMV k (with the stop test after update k-1 and direction k), update k, ...; SpMV
    // maxit only ends the solve
    // direction k is read from pb[0] at k = 0 (cg_init's p = z), else formed from direction k-1
    // in pb[(k-1) & 1] and written to pb[k & 1]
    auto spmv = [&](int k) {
      hipLaunchKernelGGL(cg_spmv<true>, g2, blk, 0, s, g, t->d_A, t->d_z, pb[k == 0 ? 0 : (k - 1) & 1], pb[k & 1],
                         t->d_q, t->d_part, t->d_S, k);
    };
    spmv(0);
    for (int k = 0; k < maxit;) {
      for (int c = 0; c < CHUNK && k < maxit; ++c, ++k) {
        // time the first iteration of each chunk (later ones may exit early once converged)
        if (c == 0) tissue_timed_begin(t, 2);
        hipLaunchKernelGGL(cg_update, g1, blk, 0, s, g, pb[k & 1], t->d_q, t->d_dinv, x, t->d_r, t->d_z,
                           t->d_part, t->d_S, k);
        if (c == 0) { tissue_timed_end(t, 2); tissue_timed_begin(t, 1); }
        spmv(k + 1);
        if (c == 0) tissue_timed_end(t, 1);
      }
      TCHK(t, hipGetLastError());
      TCHK(t, hipMemcpyAsync(&h, t->d_S, sizeof(Scal), hipMemcpyDeviceToHost, s));
      TCHK(t, hipStreamSynchronize(s));
      tissue_timed_collect(t, 1);
      tissue_timed_collect(t, 2);
      if (h.done) break;
    }
  }
  if (iters) *iters = h.iters;
  return 0;
}

static int tissue_upload_params(cep_tissue *t);

static int tissue_react_launch(cep_tissue *t, double h)
{
  cep_batch *b = t->cells;
  // the parameter table of the nodes (built by every parameter change; here if none came yet)
  if (t->nclass == 0)
    if (int rc = tissue_upload_params(t)) return rc;
  const int blocks = (t->nn + 63) / 64;
  tissue_timed_begin(t, 0);
  if (t->model == CEP_MODEL_ORD2011) {
    // paced nodes are stimulated while the beat time lies in [start, start + duration]
    // (heart.cpp:287-297); unpaced nodes carry a zero amplitude
    const double tb = t->t - floor(t->t / t->stim_bcl) * t->stim_bcl;
    const int on = (!t->s1_stopped && tb >= t->stim_start && tb <= t->stim_start + t->stim_dur) ? 1 : 0;
    if (b->prec == CEP_PREC_FP64)
      hipLaunchKernelGGL(cep::tis::tissue_react11<double>, dim3(blocks), dim3(64), 0, b->stream, t->nn, b->pitch, t->t,
                         h, on, b->d_y, t->d_ptab, t->d_pcls, t->nclass);
    else
      hipLaunchKernelGGL(cep::tis::tissue_react11<float>, dim3(blocks), dim3(64), 0, b->stream, t->nn, b->pitch, t->t,
                         h, on, b->d_y, t->d_ptab, t->d_pcls, t->nclass);
    tissue_timed_end(t, 0);
    TCHK(t, hipGetLastError());
    return 0;
  }
  if (b->prec == CEP_PREC_FP64)
    hipLaunchKernelGGL(cep::tis::tissue_react<double>, dim3(blocks), dim3(64), 0, b->stream, t->nn, b->pitch, t->t, h,
                       b->d_y, t->d_ptab, t->d_pcls, t->nclass);
  else
    hipLaunchKernelGGL(cep::tis::tissue_react<float>, dim3(blocks), dim3(64), 0, b->stream, t->nn, b->pitch, t->t, h,
                       b->d_y, t->d_ptab, t->d_pcls, t->nclass);
  tissue_timed_end(t, 0);
  TCHK(t, hipGetLastError());
  return 0;
}

// node parameters: the cell parameters with the stimulus kept only on the pacing nodes
static int tissue_upload_params(cep_tissue *t)
{
  cep_batch *b = t->cells;
  const size_t P = b->pitch;
  std::vector<double> prm(P * b->np, 0.0);
  std::vector<char> paced(t->nn, 0);
  if (!t->s1_stopped)
    for (int k : t->stim_nodes) paced[k] = 1;
  double p[64];
  for (int i = 0; i < t->nn; ++i) {
    if (t->model == CEP_MODEL_ORD2011) {
      cep::host::o11_params(b->o11_in[i], b->dutta, paced[i] ? t->stim_scale : 0.0, p);
      p[cep::ord11::P_BCL] = t->stim_bcl;
      for (int k = 0; k < b->np; ++k) prm[k * P + i] = p[k];
      continue;
    }
    cep::host::ord_params(b->ord_in[i], paced[i] ? t->stim_scale : 0.0, p);
    p[cep::ord::OP_BCL] = t->stim_bcl;
    p[cep::ord::OP_STIM_START] = t->stim_start;
    p[cep::ord::OP_STIM_DUR] = t->stim_dur;
    for (int k = 0; k < b->np; ++k) prm[k * P + i] = p[k];
  }
  TCHK(t, hipMemcpyAsync(b->d_prm, prm.data(), prm.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  // the distinct rows (bitwise) and each node's row: the reaction kernels read these
  std::map<std::vector<double>, int> rows;
  std::vector<int> cls(t->nn);
  std::vector<double> row(b->np);
  for (int i = 0; i < t->nn; ++i) {
    for (int k = 0; k < b->np; ++k) row[k] = prm[k * P + i];
    auto it = rows.emplace(row, (int)rows.size()).first;
    cls[i] = it->second;
  }
  const int K = (int)rows.size();
  std::vector<double> tab((size_t)K * b->np);
  for (const auto &r : rows)
    for (int k = 0; k < b->np; ++k) tab[(size_t)k * K + r.second] = r.first[k];
  if (K > t->ptab_cap) {
    if (talloc(t, &t->d_ptab, (size_t)K * b->np)) return CEP_E_HIP;
    t->ptab_cap = K;
  }
  if (!t->d_pcls && talloc(t, &t->d_pcls, (size_t)t->nn)) return CEP_E_HIP;
  TCHK(t, hipMemcpyAsync(t->d_ptab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  TCHK(t, hipMemcpyAsync(t->d_pcls, cls.data(), cls.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  t->nclass = K;
  TCHK(t, hipStreamSynchronize(b->stream));
  return 0;
}

extern "C" {

cep_tissue *cep_tissue_create(int nx, int ny, double hx, double hy, int precision)
{
  return cep_tissue_create_model(nx, ny, hx, hy, CEP_MODEL_ORD_CIPA, precision);
}

cep_tissue *cep_tissue_create_model(int nx, int ny, double hx, double hy, int model, int precision)
{
  if (nx < 1 || ny < 1 || !(hx > 0) || !(hy > 0)) return nullptr;
  if (model != CEP_MODEL_ORD_CIPA && model != CEP_MODEL_ORD2011) return nullptr;
  const long long nn = (long long)(nx + 1) * (ny + 1);
  if (nn > 50000000LL) return nullptr;
  cep_tissue *t = new cep_tissue();
  t->nx = nx; t->ny = ny; t->nxn = nx + 1; t->nyn = ny + 1; t->nn = (int)nn; t->hx = hx; t->hy = hy;
  t->model = model;
  t->cells = batch_create(model, t->nn, precision, false);   // no BDF workspace
  if (!t->cells) { delete t; return nullptr; }
  t->sxx.assign((size_t)nx * ny, 0.0);
  t->sxy.assign((size_t)nx * ny, 0.0);
  t->syy.assign((size_t)nx * ny, 0.0);
  t->cp = (t->nxn + cep::tis::TX - 1) / cep::tis::TX * cep::tis::TX;
  const int nblk = tissue_grid(t).maxblk;
  int rc = 0;
  rc |= talloc(t, &t->d_A, (size_t)9 * t->cp * t->nyn);
  rc |= talloc(t, &t->d_B, (size_t)9 * t->cp * t->nyn);
  rc |= talloc(t, &t->d_p2, t->nn);
  rc |= talloc(t, &t->d_dinv, t->nn);
  rc |= talloc(t, &t->d_b, t->nn);
  rc |= talloc(t, &t->d_r, t->nn);
  rc |= talloc(t, &t->d_z, t->nn);
  rc |= talloc(t, &t->d_p, t->nn);
  rc |= talloc(t, &t->d_q, t->nn);
  rc |= talloc(t, &t->d_part, (size_t)cep::tis::P_N * nblk);
  rc |= talloc(t, &t->d_S, 1);
  t->ev.resize(6);
  for (auto &e : t->ev) rc |= hipEventCreate(&e) != hipSuccess;
  if (rc || hipStreamSynchronize(t->cells->stream) != hipSuccess) {
    cep_tissue_destroy(t);
    return nullptr;
  }
  return t;
}

cep_tissue *cep_tissue_create_mesh(const cep_mesh *m, int model, int precision)
{
  if (!m) return nullptr;
  // the element types the reference's FEM assembles (fem.cpp:830-860): square, hex, tet
  int kind;
  if (m->npe == 4 && m->dim == 2) kind = EK_QUAD4;
  else if (m->npe == 8 && m->dim == 3) kind = EK_HEX8;
  else if (m->npe == 4 && m->dim == 3) kind = EK_TET4;
  else return nullptr;
  if (model != CEP_MODEL_ORD_CIPA && model != CEP_MODEL_ORD2011) return nullptr;
  if (m->nn > 50000000) return nullptr;
  const int n = m->npe;
  // every element must map the reference element one-to-one: det J > 0 at its Gauss points
  {
    const ElemTable E = elem_table(kind);
    const double D[6] = { 1, 0, 0, 1, 0, 1 };
    double Me[64], Ke[64], X[8][3];
    for (int e = 0; e < m->ne; ++e) {
      for (int a = 0; a < n; ++a)
        for (int c = 0; c < 3; ++c) X[a][c] = m->xyz[(size_t)m->conn[(size_t)e * n + a] * 3 + c];
      if (!(elem_matrices(E, X, D, Me, Ke) > 0.0)) return nullptr;
    }
  }
  cep_tissue *t = new cep_tissue();
  t->mesh = true;
  t->ekind = kind; t->npe = n; t->dim = kind == EK_QUAD4 ? 2 : 3;
  t->nn = m->nn; t->ne = m->ne; t->model = model;
  t->xyz = m->xyz;
  t->conn = m->conn;
  // CSR pattern: every pair of nodes that share an element, columns ascending
  std::vector<std::vector<int>> cols(t->nn);
  for (int e = 0; e < t->ne; ++e)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) cols[t->conn[(size_t)e * n + j]].push_back(t->conn[(size_t)e * n + k]);
  t->rp.assign(t->nn + 1, 0);
  for (int i = 0; i < t->nn; ++i) {
    auto &c = cols[i];
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    if (c.empty()) c.push_back(i);                     // a node in no element: identity row
    t->rp[i + 1] = t->rp[i] + (int)c.size();
  }
  t->nnz = t->rp[t->nn];
  t->ci.reserve(t->nnz);
  for (auto &c : cols) t->ci.insert(t->ci.end(), c.begin(), c.end());
  t->dten.assign((size_t)t->ne * 6, 0.0);
  t->cells = batch_create(model, t->nn, precision, false);
  if (!t->cells) { delete t; return nullptr; }
  const int nblk = tissue_grid(t).maxblk;
  int rc = 0;
  rc |= talloc(t, &t->d_A, (size_t)t->nnz);
  rc |= talloc(t, &t->d_B, (size_t)t->nnz);
  rc |= talloc(t, &t->d_rp, (size_t)t->nn + 1);
  rc |= talloc(t, &t->d_ci, (size_t)t->nnz);
  rc |= talloc(t, &t->d_dinv, t->nn);
  rc |= talloc(t, &t->d_b, t->nn);
  rc |= talloc(t, &t->d_r, t->nn);
  rc |= talloc(t, &t->d_z, t->nn);
  rc |= talloc(t, &t->d_p, t->nn);
  rc |= talloc(t, &t->d_q, t->nn);
  rc |= talloc(t, &t->d_part, (size_t)cep::tis::P_N * nblk);
  rc |= talloc(t, &t->d_S, 1);
  t->ev.resize(6);
  for (auto &e : t->ev) rc |= hipEventCreate(&e) != hipSuccess;
  if (!rc) {
    hipStream_t st = t->cells->stream;
    rc |= hipMemcpyAsync(t->d_rp, t->rp.data(), t->rp.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess;
    rc |= hipMemcpyAsync(t->d_ci, t->ci.data(), t->ci.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess;
  }
  if (rc || hipStreamSynchronize(t->cells->stream) != hipSuccess) {
    cep_tissue_destroy(t);
    return nullptr;
  }
  return t;
}

int cep_tissue_write_checkpoint(const cep_tissue *t, const char *path)
{
  if (!t || !path) return CEP_E_ARG;
  cep_batch *b = t->cells;
  std::vector<double> y((size_t)t->nn * b->ns);
  if (int rc = cep_tissue_get_states(t, y.data())) return rc;
  FILE *f = fopen(path, "w");
  if (!f) return tfail(const_cast<cep_tissue *>(t), CEP_E_ARG, std::string("cannot write ") + path);
  fprintf(f, "%f \n", t->t);
  for (int i = 0; i < t->nn; ++i) {
    for (int j = 0; j < b->ns; ++j) fprintf(f, "%f ", y[(size_t)i * b->ns + j]);
    fprintf(f, "\n");
  }
  fclose(f);
  return 0;
}

int cep_tissue_read_checkpoint(cep_tissue *t, const char *path)
{
  if (!t || !path) return CEP_E_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return tfail(t, CEP_E_ARG, std::string("cannot read ") + path);
  cep_batch *b = t->cells;
  double time = 0.0;
  std::vector<double> y((size_t)t->nn * b->ns);
  bool ok = fscanf(f, "%lf", &time) == 1;
  for (size_t k = 0; ok && k < y.size(); ++k) ok = fscanf(f, "%lf", &y[k]) == 1;
  fclose(f);
  if (!ok) return tfail(t, CEP_E_ARG, std::string("short or malformed checkpoint ") + path);
  if (int rc = cep_tissue_set_states(t, y.data())) return rc;
  t->t = time;
  return 0;
}

void cep_tissue_destroy(cep_tissue *t)
{
  if (!t) return;
  if (t->cells) (void)hipStreamSynchronize(t->cells->stream);
  for (void *p : t->allocs) (void)hipFree(p);
  for (auto &e : t->ev) if (e) (void)hipEventDestroy(e);
  if (t->cells) cep_batch_destroy(t->cells);
  delete t;
}

const char *cep_tissue_last_error(const cep_tissue *t)
{
  if (!t) return "null tissue";
  if (!t->err.empty()) return t->err.c_str();
  return t->cells ? cep_last_error(t->cells) : "";
}

int cep_tissue_set_cells(cep_tissue *t, const double *celltype, const double *conc, const double *herg,
                         int herg_stride, const double *gscale, int gscale_stride)
{
  if (!t) return CEP_E_ARG;
  if (int rc = cep_set_cells(t->cells, 0, t->nn, celltype, conc, herg, herg_stride, gscale, gscale_stride, nullptr))
    return tfail(t, rc, cep_last_error(t->cells));
  // initial states of every node (initConsts), D = the node's concentration
  cep_batch *b = t->cells;
  const size_t P = b->pitch;
  std::vector<double> y(P * b->ns, 0.0);
  for (int i = 0; i < t->nn; ++i) {
    if (t->model == CEP_MODEL_ORD2011) {
      for (int k = 0; k < cep::ord11::NS; ++k) y[k * P + i] = cep::host::kO11Init[k];
      continue;
    }
    for (int k = 0; k < 49; ++k) y[k * P + i] = cep::host::kOrdInitStates[k];
    y[43 * P + i] = b->ord_in[i].conc;
  }
  TCHK(t, hipMemcpyAsync(b->d_y, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  TCHK(t, hipStreamSynchronize(b->stream));
  t->t = 0.0;
  return tissue_upload_params(t);
}

// a symmetric 3x3 tensor (xx, xy, xz, yy, yz, zz) is positive semi-definite when every
// principal minor is non-negative
static bool tensor_psd(const double *D)
{
  const double xx = D[0], xy = D[1], xz = D[2], yy = D[3], yz = D[4], zz = D[5];
  if (!(xx >= 0) || !(yy >= 0) || !(zz >= 0)) return false;
  const double tol = 1e-12 * (xx + yy + zz) * (xx + yy + zz);
  if (xx * yy - xy * xy < -tol || xx * zz - xz * xz < -tol || yy * zz - yz * yz < -tol) return false;
  const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
  return det >= -tol * (xx + yy + zz);
}

int cep_tissue_set_conductivity(cep_tissue *t, const double *sxx, const double *sxy, const double *syy)
{
  if (!t || !sxx || !syy) return tfail(t, CEP_E_ARG, "conductivity arrays");
  if (t->mesh && t->dim == 3) return tfail(t, CEP_E_ARG, "3-D mesh: give the tensor with cep_tissue_set_conductivity_tensor");
  if (t->mesh) {
    std::vector<double> d((size_t)t->ne * 3);
    for (int e = 0; e < t->ne; ++e) { d[e * 3] = sxx[e]; d[e * 3 + 1] = sxy ? sxy[e] : 0.0; d[e * 3 + 2] = syy[e]; }
    return cep_tissue_set_conductivity_tensor(t, d.data(), 3);
  }
  const size_t ne = (size_t)t->nx * t->ny;
  for (size_t e = 0; e < ne; ++e) {
    t->sxx[e] = sxx[e];
    t->sxy[e] = sxy ? sxy[e] : 0.0;
    t->syy[e] = syy[e];
    if (!(sxx[e] >= 0) || !(syy[e] >= 0) || t->sxy[e] * t->sxy[e] > sxx[e] * syy[e])
      return tfail(t, CEP_E_ARG, "diffusion tensor not positive semi-definite");
  }
  t->dt_built = -1;
  return 0;
}

int cep_tissue_set_conductivity_tensor(cep_tissue *t, const double *d, int ncomp)
{
  if (!t || !d) return tfail(t, CEP_E_ARG, "conductivity array");
  const int dim = t->mesh ? t->dim : 2;
  if (!(ncomp == 1 || (dim == 2 && ncomp == 3) || (dim == 3 && ncomp == 6)))
    return tfail(t, CEP_E_ARG, "conductivity components: 1 (isotropic), 3 (2-D) or 6 (3-D)");
  const size_t ne = t->mesh ? (size_t)t->ne : (size_t)t->nx * t->ny;
  std::vector<double> D(ne * 6, 0.0);
  for (size_t e = 0; e < ne; ++e) {
    const double *v = d + e * ncomp;
    double *o = &D[e * 6];
    if (ncomp == 1) { o[0] = v[0]; o[3] = v[0]; o[5] = dim == 3 ? v[0] : 0.0; }
    else if (ncomp == 3) { o[0] = v[0]; o[1] = v[1]; o[3] = v[2]; }
    else for (int k = 0; k < 6; ++k) o[k] = v[k];
    if (!tensor_psd(o)) return tfail(t, CEP_E_ARG, "diffusion tensor not positive semi-definite");
  }
  if (t->mesh) {
    t->dten.swap(D);
  } else {
    for (size_t e = 0; e < ne; ++e) { t->sxx[e] = D[e * 6]; t->sxy[e] = D[e * 6 + 1]; t->syy[e] = D[e * 6 + 3]; }
  }
  t->dt_built = -1;
  return 0;
}

int cep_tissue_set_stimulus(cep_tissue *t, const int *nodes, int count, double amp_scale, double bcl,
                            double start, double duration)
{
  if (!t || count < 0 || (count && !nodes) || !(bcl > 0) || !(duration >= 0)) return tfail(t, CEP_E_ARG, "stimulus");
  t->stim_nodes.clear();
  for (int k = 0; k < count; ++k) {
    if (nodes[k] < 0 || nodes[k] >= t->nn) return tfail(t, CEP_E_ARG, "stimulus node out of range");
    t->stim_nodes.push_back(nodes[k]);
  }
  t->stim_scale = amp_scale; t->stim_bcl = bcl; t->stim_start = start; t->stim_dur = duration;
  return tissue_upload_params(t);
}

int cep_tissue_set_s1s2(cep_tissue *t, int npace1, const int *s2_nodes, int count, int target_node)
{
  if (!t || count < 0 || (count && !s2_nodes) || target_node < 0 || target_node >= t->nn)
    return tfail(t, CEP_E_ARG, "S1-S2 arguments");
  cep_batch *b = t->cells;
  const int ns = b->ns;
  std::vector<double> init((size_t)(count > 0 ? count : 1) * ns, 0.0);
  for (int k = 0; k < count; ++k) {
    const int n = s2_nodes[k];
    if (n < 0 || n >= t->nn) return tfail(t, CEP_E_ARG, "S2 node out of range");
    double *y = &init[(size_t)k * ns];
    if (t->model == CEP_MODEL_ORD2011) {
      for (int j = 0; j < cep::ord11::NS; ++j) y[j] = cep::host::kO11Init[j];
    } else {
      for (int j = 0; j < 49; ++j) y[j] = cep::host::kOrdInitStates[j];
      y[43] = b->ord_in[n].conc;
    }
  }
  if (!t->d_s2) {
    int rc = talloc(t, &t->d_s2, 1);
    if (rc) return rc;
  }
  if (count > t->s2_count || !t->d_s2nodes) {
    int rc = talloc(t, &t->d_s2nodes, (size_t)(count > 0 ? count : 1));
    rc |= talloc(t, &t->d_s2init, init.size());
    if (rc) return rc;
  }
  hipStream_t st = b->stream;
  const cep::tis::S2State zero{ 0, 0, 0, 0, -1.0 };
  TCHK(t, hipMemcpyAsync(t->d_s2, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
  if (count) {
    TCHK(t, hipMemcpyAsync(t->d_s2nodes, s2_nodes, (size_t)count * sizeof(int), hipMemcpyHostToDevice, st));
    TCHK(t, hipMemcpyAsync(t->d_s2init, init.data(), init.size() * sizeof(double), hipMemcpyHostToDevice, st));
  }
  TCHK(t, hipStreamSynchronize(st));
  t->npace1 = npace1;
  t->s2_count = count;
  t->s2_target = target_node;
  if (t->s1_stopped) {                         // a new protocol: pacing on again
    t->s1_stopped = false;
    return tissue_upload_params(t);
  }
  return 0;
}

int cep_tissue_s2_time(const cep_tissue *t, double *time)
{
  if (!t || !time) return CEP_E_ARG;
  *time = -1.0;
  if (!t->d_s2) return 0;
  cep::tis::S2State h{};
  TCHK(const_cast<cep_tissue *>(t), hipMemcpy(&h, t->d_s2, sizeof(h), hipMemcpyDeviceToHost));
  *time = h.done ? h.time : -1.0;
  return 0;
}

int cep_tissue_step(cep_tissue *t, int nsteps, double dt, int split, double cg_rtol, int cg_maxit,
                    long long *cg_iters)
{
  if (!t || nsteps < 0 || !(dt > 0) || (split != CEP_SPLIT_GODUNOV && split != CEP_SPLIT_STRANG) ||
      !(cg_rtol > 0) || cg_maxit < 1)
    return tfail(t, CEP_E_ARG, "tissue step arguments");
  long long total = 0;
  double *vm = t->cells->d_y;                  // state plane 0
  for (int s = 0; s < nsteps; ++s) {
    const double t0 = t->t;
    if (t->npace1 > 0) {                       // heart.cpp:315-323, before the node loop
      const int pace = (int)(t0 / t->stim_bcl);
      if (pace >= t->npace1 && !t->s1_stopped) {
        t->s1_stopped = true;
        if (int rc = tissue_upload_params(t)) return rc;
      }
      if (pace == t->npace1 - 1 && t->s2_count > 0) {
        cep_batch *b = t->cells;
        hipLaunchKernelGGL(cep::tis::tissue_s2, dim3(1), dim3(256), 0, b->stream, t->s2_target, b->pitch, b->ns, t0,
                           b->d_y, t->d_s2, t->d_s2nodes, t->s2_count, t->d_s2init);
      }
    }
    if (split == CEP_SPLIT_GODUNOV) {
      if (int rc = tissue_react_launch(t, dt)) return rc;
    } else {
      if (int rc = tissue_react_launch(t, 0.5 * dt)) return rc;
    }
    int it = 0;
    if (int rc = tissue_solve(t, dt, vm, cg_rtol, cg_maxit, &it)) return rc;
    tissue_timed_collect(t, 0);
    total += it;
    if (split == CEP_SPLIT_STRANG) {
      // not timed: waiting for its event would add a host round trip to every step, during
      // which the device idles (the first half-step gives the reaction's timing)
      t->t = t0 + 0.5 * dt;
      if (int rc = tissue_react_launch(t, 0.5 * dt)) return rc;
    }
    t->t = t0 + dt;
  }
  TCHK(t, hipStreamSynchronize(t->cells->stream));
  if (cg_iters) *cg_iters = total;
  return 0;
}

int cep_tissue_diffuse(cep_tissue *t, double dt, const double *v, double *x, double cg_rtol, int cg_maxit, int *iters)
{
  if (!t || !v || !x || !(dt > 0) || !(cg_rtol > 0) || cg_maxit < 1) return tfail(t, CEP_E_ARG, "diffuse arguments");
  hipStream_t s = t->cells->stream;
  double *xv = nullptr;
  TCHK(t, hipMalloc(&xv, (size_t)t->nn * sizeof(double)));
  TCHK(t, hipMemcpyAsync(xv, v, (size_t)t->nn * sizeof(double), hipMemcpyHostToDevice, s));
  int rc = tissue_solve(t, dt, xv, cg_rtol, cg_maxit, iters);
  tissue_timed_collect(t, 1);
  if (!rc) {
    if (hipMemcpyAsync(x, xv, (size_t)t->nn * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = tfail(t, CEP_E_HIP, "diffuse copy back");
  }
  (void)hipFree(xv);
  return rc;
}

int cep_tissue_get_vm(const cep_tissue *t, double *out)
{
  if (!t || !out) return CEP_E_ARG;
  hipStream_t s = t->cells->stream;
  if (hipMemcpyAsync(out, t->cells->d_y, (size_t)t->nn * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return CEP_E_HIP;
  return 0;
}

int cep_tissue_get_states(const cep_tissue *t, double *out)
{
  if (!t) return CEP_E_ARG;
  return cep_get_states(t->cells, out);
}

int cep_tissue_set_states(cep_tissue *t, const double *in)
{
  if (!t || !in) return CEP_E_ARG;
  cep_batch *b = t->cells;
  const size_t P = b->pitch;
  std::vector<double> y(P * b->ns, 0.0);
  for (int i = 0; i < t->nn; ++i)
    for (int k = 0; k < b->ns; ++k) y[k * P + i] = in[(size_t)i * b->ns + k];
  TCHK(t, hipMemcpyAsync(b->d_y, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  TCHK(t, hipStreamSynchronize(b->stream));
  return 0;
}

double cep_tissue_time(const cep_tissue *t) { return t ? t->t : 0.0; }

int cep_tissue_kernel_times(cep_tissue *t, double *avg_ms, long long *n)
{
  if (!t) return CEP_E_ARG;
  for (int k = 0; k < 3; ++k) {
    if (avg_ms) avg_ms[k] = t->tcount[k] ? t->tsum[k] / t->tcount[k] : 0.0;
    if (n) n[k] = t->tcount[k];
    t->tsum[k] = 0.0;
    t->tcount[k] = 0;
  }
  return 0;
}

}  // extern "C"
