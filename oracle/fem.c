/*
 * fem.c -- ORACLE (test infrastructure only): the tissue diffusion step of fem.cpp.
 *
 * fem_q1_element: bilinear quad element matrices as Cfem::set_amat_bmat builds them
 * (fem.cpp:628-661) with the reference's 2x2 Gauss points at +-0.57735
 * (set_gpt_square, :256-270) and shape functions of set_Phi_dPhiS_square (:398-426),
 * for an axis-aligned hx x hy element and a 2x2 diffusion tensor (the reference's
 * isotropic case is sxx = syy = DE, sxy = 0).  Me is the mass term, Ke the conduction
 * term; the reference's amat is Me/dt + Ke and bmat is Me.  Node order: (0,0), (hx,0),
 * (hx,hy), (0,hy).
 * fem_q4_element: the same element on any quadrilateral (x[4], y[4] in the node order
 * above): at each Gauss point the Jacobian of the bilinear map from the node coordinates,
 * its determinant and inverse give the physical gradients, as set_amat_bmat does with
 * MultMatMat / inverseMat (:631-640) for every mesh element.
 * cg_solve_csr: CG with Jacobi preconditioning from a zero initial guess, stopped on
 * the preconditioned residual norm relative to its initial value, as PETSc's KSPCG +
 * PCJACOBI with default tolerances (rtol 1e-5) used by Cfem::assemble_Amat_Bmat
 * (:753-760) and solve_Bmat (:705-718).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void fem_q1_element(double hx, double hy, double sxx, double sxy, double syy, double *Me, double *Ke)
{
  const double g[2] = { -0.57735, 0.57735 };
  const double gs[4] = { g[0], g[1], g[1], g[0] }, gt[4] = { g[0], g[0], g[1], g[1] };
  memset(Me, 0, 16 * sizeof(double));
  memset(Ke, 0, 16 * sizeof(double));
  const double detJ = 0.25 * hx * hy;
  for (int gp = 0; gp < 4; ++gp) {
    const double s = gs[gp], t = gt[gp];
    const double phi[4] = { (1. - s) * (1. - t) / 4., (1. + s) * (1. - t) / 4.,
                            (1. + s) * (1. + t) / 4., (1. - s) * (1. + t) / 4. };
    const double dps[4] = { -(1. - t) / 4., (1. - t) / 4., (1. + t) / 4., -(1. + t) / 4. };
    const double dpt[4] = { -(1. - s) / 4., -(1. + s) / 4., (1. + s) / 4., (1. - s) / 4. };
    double dx[4], dy[4];
    for (int a = 0; a < 4; ++a) { dx[a] = dps[a] * (2.0 / hx); dy[a] = dpt[a] * (2.0 / hy); }
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        Me[j * 4 + k] += phi[j] * phi[k] * detJ;
        Ke[j * 4 + k] += (dx[j] * (sxx * dx[k] + sxy * dy[k]) + dy[j] * (sxy * dx[k] + syy * dy[k])) * detJ;
      }
  }
}

void fem_q4_element(const double *xe, const double *ye, double sxx, double sxy, double syy, double *Me, double *Ke)
{
  const double g[2] = { -0.57735, 0.57735 };
  const double gs[4] = { g[0], g[1], g[1], g[0] }, gt[4] = { g[0], g[0], g[1], g[1] };
  memset(Me, 0, 16 * sizeof(double));
  memset(Ke, 0, 16 * sizeof(double));
  for (int gp = 0; gp < 4; ++gp) {
    const double s = gs[gp], t = gt[gp];
    const double phi[4] = { (1. - s) * (1. - t) / 4., (1. + s) * (1. - t) / 4.,
                            (1. + s) * (1. + t) / 4., (1. - s) * (1. + t) / 4. };
    const double dps[4] = { -(1. - t) / 4., (1. - t) / 4., (1. + t) / 4., -(1. + t) / 4. };
    const double dpt[4] = { -(1. - s) / 4., -(1. + s) / 4., (1. + s) / 4., (1. - s) / 4. };
    /* Jacobian (rows: d/ds, d/dt; columns: x, y) = dPhiS^T * xyele */
    double J[2][2] = { { 0, 0 }, { 0, 0 } };
    for (int a = 0; a < 4; ++a) {
      J[0][0] += dps[a] * xe[a]; J[0][1] += dps[a] * ye[a];
      J[1][0] += dpt[a] * xe[a]; J[1][1] += dpt[a] * ye[a];
    }
    const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double inv[2][2] = { { J[1][1] / det, -J[0][1] / det }, { -J[1][0] / det, J[0][0] / det } };
    double dx[4], dy[4];
    for (int a = 0; a < 4; ++a) {
      dx[a] = inv[0][0] * dps[a] + inv[0][1] * dpt[a];
      dy[a] = inv[1][0] * dps[a] + inv[1][1] * dpt[a];
    }
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        Me[j * 4 + k] += phi[j] * phi[k] * det;
        Ke[j * 4 + k] += (dx[j] * (sxx * dx[k] + sxy * dy[k]) + dy[j] * (sxy * dx[k] + syy * dy[k])) * det;
      }
  }
}

int cg_solve_csr(int n, const int *rowptr, const int *col, const double *val,
                 const double *b, double *x, double rtol, int maxit, int jacobi)
{
  double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
  double *dinv = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) {
    dinv[i] = 1.0;
    if (jacobi)
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) if (col[k] == i) dinv[i] = 1.0 / val[k];
    x[i] = 0.0;
    r[i] = b[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
  }
  double rz = 0.0, zz = 0.0;
  for (int i = 0; i < n; ++i) { rz += r[i] * z[i]; zz += z[i] * z[i]; }
  const double stop = rtol * sqrt(zz);
  int it = 0;
  if (sqrt(zz) <= 0.0) goto done;
  for (it = 1; it <= maxit; ++it) {
    double pq = 0.0;
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * p[col[k]];
      q[i] = s;
      pq += p[i] * s;
    }
    const double alpha = rz / pq;
    double rz1 = 0.0;
    zz = 0.0;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
      rz1 += r[i] * z[i];
      zz += z[i] * z[i];
    }
    if (sqrt(zz) <= stop) break;
    const double beta = rz1 / rz;
    rz = rz1;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
done:
  free(r); free(z); free(p); free(q); free(dinv);
  return it;
}
