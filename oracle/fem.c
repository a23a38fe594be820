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
 * fem_hex8_element: the trilinear hexahedron (C3D8I, "hex") of set_gpt_hex (:231-248, points
 * +-0.57735 in the order of its ix, iy, iz loops) and set_Phi_dPhiS_hex (:427-479, nodes
 * (-,-,-), (+,-,-), (+,+,-), (-,+,-), then the same at z = +1); 3x3 Jacobian, det3D and
 * inverseMat (:549-589).  D is the tensor (xx, xy, xz, yy, yz, zz); Me and Ke are 8x8.
 * fem_tet4_element: the linear tetrahedron (C3D4, "tet") of set_gpt_tet (:209-229, points
 * (5 -+ sqrt 5)/20) and set_Phi_dPhiS_tet (:342-374, phi = 1-s-t-z, s, t, z), 4x4.  The
 * reference adds its four points with weight 1; the weight here is 1/24, the volume of the
 * reference tetrahedron over four: both A and B of the reference are 24 times these, which
 * leaves A x = B v unchanged.
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

/* the physical gradients G (n x 3) of n shape functions with reference derivatives dS
   (n x 3) through the 3x3 Jacobian J_rc = sum_a dS[a][r] X[a][c]; returns det J */
static double grad3(int n, const double *X, const double (*dS)[3], double (*G)[3])
{
  double J[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      J[r][c] = 0.0;
      for (int a = 0; a < n; ++a) J[r][c] += dS[a][r] * X[a * 3 + c];
    }
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double inv[3][3];
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  inv[1][0] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  inv[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  inv[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  /* dphi/dx_c = sum_r (J^-1)_{c r} dphi/ds_r */
  for (int a = 0; a < n; ++a)
    for (int c = 0; c < 3; ++c) G[a][c] = inv[c][0] * dS[a][0] + inv[c][1] * dS[a][1] + inv[c][2] * dS[a][2];
  return det;
}

static void add_point(int n, const double *phi, const double (*G)[3], const double *D, double wdet, double *Me,
                      double *Ke)
{
  const double Dm[3][3] = { { D[0], D[1], D[2] }, { D[1], D[3], D[4] }, { D[2], D[4], D[5] } };
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      double kk = 0.0;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) kk += G[j][r] * Dm[r][c] * G[k][c];
      Me[j * n + k] += phi[j] * phi[k] * wdet;
      Ke[j * n + k] += kk * wdet;
    }
}

void fem_hex8_element(const double *X, const double *D, double *Me, double *Ke)
{
  const double g[2] = { -0.57735, 0.57735 };
  const double ns[8] = { -1, 1, 1, -1, -1, 1, 1, -1 }, nt[8] = { -1, -1, 1, 1, -1, -1, 1, 1 };
  const double nz[8] = { -1, -1, -1, -1, 1, 1, 1, 1 };
  memset(Me, 0, 64 * sizeof(double));
  memset(Ke, 0, 64 * sizeof(double));
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy)
      for (int iz = 0; iz < 2; ++iz) {
        const double s = g[ix], t = g[iy], z = g[iz];
        double phi[8], dS[8][3], G[8][3];
        for (int a = 0; a < 8; ++a) {
          phi[a] = (1 + ns[a] * s) * (1 + nt[a] * t) * (1 + nz[a] * z) / 8.;
          dS[a][0] = ns[a] * (1 + nt[a] * t) * (1 + nz[a] * z) / 8.;
          dS[a][1] = nt[a] * (1 + ns[a] * s) * (1 + nz[a] * z) / 8.;
          dS[a][2] = nz[a] * (1 + ns[a] * s) * (1 + nt[a] * t) / 8.;
        }
        const double det = grad3(8, X, (const double(*)[3])dS, G);
        add_point(8, phi, (const double(*)[3])G, D, det, Me, Ke);
      }
}

void fem_tet4_element(const double *X, const double *D, double *Me, double *Ke)
{
  const double aa = (5. - sqrt(5.)) / 20., bb = (5. + 3. * sqrt(5.)) / 20.;
  const double P[4][3] = { { aa, aa, aa }, { aa, aa, bb }, { aa, bb, aa }, { bb, aa, aa } };
  const double dS[4][3] = { { -1, -1, -1 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
  memset(Me, 0, 16 * sizeof(double));
  memset(Ke, 0, 16 * sizeof(double));
  for (int q = 0; q < 4; ++q) {
    const double phi[4] = { 1. - P[q][0] - P[q][1] - P[q][2], P[q][0], P[q][1], P[q][2] };
    double G[4][3];
    const double det = grad3(4, X, dS, G);
    add_point(4, phi, (const double(*)[3])G, D, det / 24.0, Me, Ke);
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
