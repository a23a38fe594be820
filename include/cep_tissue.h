/*
 * cep_tissue.h -- C ABI of the 2-D monodomain tissue step (libcepgpu.so).
 *
 * Replaces the tissue driver's per-step work of the reference:
 *   Cheart::cellular_reaction (heart.cpp:274-430): per node, computeRates + an explicit
 *     step at fixed dt, S1 stimulus on the pacing nodes;
 *   Cfem::set_amat_bmat / assemble_Amat_Bmat / solve_Bmat (fem.cpp:628-661,721-763,705-718):
 *     bilinear (Q1) elements with 2x2 Gauss points, A = M/dt + K and B = M/dt (consistent
 *     mass), b = B V^n, A V^{n+1} = b solved by Jacobi-preconditioned CG (PETSc KSPCG +
 *     PCJACOBI, relative tolerance on the preconditioned residual, zero initial guess).
 * The mesh here is a structured nx x ny grid of hx x hy elements (nodes (nx+1) x (ny+1),
 * node (ix, iy) has index iy*(nx+1) + ix) with a per-element diffusion tensor, so the
 * global matrix is a 9-point stencil per node stored as nine coefficient planes.
 * Each node carries an ORd-CiPA cell (the population batch's cell parameters and states), or an
 * ORd 2011 cell (cep_tissue_create_model).
 */
#ifndef CEP_TISSUE_H
#define CEP_TISSUE_H

#include "cep_cellsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cep_tissue cep_tissue;

/* Grid of nx x ny elements of size hx x hy (mm); every node an ORd-CiPA cell with the
   given precision of its ionic model (CEP_PREC_*).  Diffusion is always FP64. */
cep_tissue *cep_tissue_create(int nx, int ny, double hx, double hy, int precision);
/* The same with the node model chosen: CEP_MODEL_ORD_CIPA, or CEP_MODEL_ORD2011, the model the
   reference's tissue driver uses (heart.h:10-14) with its own reaction scheme: computeRates
   with the TISSUE stimulus (isS1 ? amp : 0) and solveAnalytical(dt) (heart.cpp:341-368);
   paced nodes are stimulated while the beat time lies in [start, start + duration]
   (heart.cpp:287-297).  ORd 2011 cells take 14-value IC50 rows and no population scales. */
cep_tissue *cep_tissue_create_model(int nx, int ny, double hx, double hy, int model, int precision);
void cep_tissue_destroy(cep_tissue *t);
const char *cep_tissue_last_error(const cep_tissue *t);

/* Cell parameters of every node (same meaning as cep_set_cells; arrays of nnodes rows,
   null = defaults) and the initial states (initConsts); nodes are not stimulated until
   cep_tissue_set_stimulus names them. */
int cep_tissue_set_cells(cep_tissue *t, const double *celltype, const double *conc, const double *herg,
                         int herg_stride, const double *gscale, int gscale_stride);
/* Diffusion tensor per element (nx*ny each, mm^2/ms): the reference's DE*I for the
   isotropic case (fem.cpp:787-828), or the fibre-rotated tensor. */
int cep_tissue_set_conductivity(cep_tissue *t, const double *sxx, const double *sxy, const double *syy);
/* The same for every tissue kind, with ncomp values per element: 1 = the reference's
   isotropic DE (Cfem::make_diffusion_tensor, fem.cpp:162-170, D = DE*I), 3 = a 2-D tensor
   (xx, xy, yy), 6 = a 3-D tensor of a hexahedral or tetrahedral mesh (xx, xy, xz, yy, yz,
   zz).  Tensors must be positive semi-definite. */
int cep_tissue_set_conductivity_tensor(cep_tissue *t, const double *d, int ncomp);
/* S1 pacing: the listed nodes get the model's stimulus (amplitude scale, period bcl, start
   and duration in ms; the reference's isS1 nodes, heart.cpp:287-318). */
int cep_tissue_set_stimulus(cep_tissue *t, const int *nodes, int count, double amp_scale, double bcl,
                            double start, double duration);
/* S1-S2 protocol of the reference driver (heart.cpp:315-323, Cheart::apply_s2 :475-501):
   S1 pacing stops once int(t / bcl) >= npace1 (npace1 <= 0: never); during the last S1
   pace, once the target node's Vm has risen above -45 mV and fallen below it again, the
   listed S2 nodes are reset to their initial states (once).  Call after cep_tissue_set_cells
   (the initial states are the nodes' own) and cep_tissue_set_stimulus (bcl). */
int cep_tissue_set_s1s2(cep_tissue *t, int npace1, const int *s2_nodes, int count, int target_node);
/* time (ms) at which the S2 nodes were reset, or a negative value if not yet */
int cep_tissue_s2_time(const cep_tissue *t, double *time);

#define CEP_SPLIT_GODUNOV 0    /* reaction(dt), then diffusion(dt): the reference's order */
#define CEP_SPLIT_STRANG 1     /* reaction(dt/2), diffusion(dt), reaction(dt/2) */

/* nsteps time steps of size dt (ms).  cg_rtol: relative tolerance of the CG stop test
   (1e-5 = PETSc's default); cg_maxit per solve.  Returns 0 or a CEP_E* code; total CG
   iterations of the call in *cg_iters (may be null). */
int cep_tissue_step(cep_tissue *t, int nsteps, double dt, int split, double cg_rtol, int cg_maxit,
                    long long *cg_iters);

/* Only the diffusion solve A x = B v (for parity tests): v, x host arrays of nnodes. */
int cep_tissue_diffuse(cep_tissue *t, double dt, const double *v, double *x, double cg_rtol, int cg_maxit,
                       int *iters);

int cep_tissue_get_vm(const cep_tissue *t, double *out);              /* nnodes */
int cep_tissue_get_states(const cep_tissue *t, double *out);          /* nnodes x nstates (52 / 41) */
int cep_tissue_set_states(cep_tissue *t, const double *in);           /* nnodes x nstates (52 / 41) */
/* simulated time (ms) */
double cep_tissue_time(const cep_tissue *t);
/* Average kernel time (ms) per call of the reaction kernel, the SpMV and the CG update
   kernels since the last query, timed with HIP events; launches counted in n[3]. */
int cep_tissue_kernel_times(cep_tissue *t, double *avg_ms, long long *n);

#ifdef __cplusplus
}
#endif
#endif
