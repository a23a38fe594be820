/*
 * cep_mesh.h -- tissue meshes, pacing-site files and state checkpoints of the reference's
 * tissue driver (libcepgpu.so, host code), and the tissue built on such a mesh.
 *
 *   mesh file      Cmesh::loadInp (mesh.cpp:59-183): a "CML" line followed by
 *                  "all_nodes all_elements target0 target1 target2", "*NODE" and one
 *                  "id, x, y, z" line per node, a "**HWCOLOR" line, "*ELEMENT, TYPE=..." and one
 *                  "id, n1, ..., nk" line per element (1-based), then an optional "CELLTYPE"
 *                  block (one integer per node) and "eof"
 *   pacing file    Cheart::loadPacingsites (heart.cpp:966-998): a count, then that many
 *                  1-based node numbers
 *   checkpoint     Cheart::writeoutput (heart.cpp:942-962): "%f \n" with the time, then one
 *                  line per node with every state as "%f "
 * Element types: the reference reads S2/T3D2 (fibres), S3, S4, C3D4 and C3D8I.  The tissue
 * solver here takes S4 (bilinear quadrilaterals, the reference's "square" mesh type); the
 * reader accepts every type.  All functions return 0 or a negative CEP_E_* code.
 */
#ifndef CEP_MESH_H
#define CEP_MESH_H

#include "cep_tissue.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cep_mesh cep_mesh;

/* Read a mesh file; null on a missing file or a malformed section. */
cep_mesh *cep_mesh_read_inp(const char *path);
/* A mesh from arrays: xyz (nnodes x 3), conn (nelem x nodes_per_elem, 0-based), celltype
   (nnodes, may be null).  nodes_per_elem 4 with dimension 2 is an S4 mesh. */
cep_mesh *cep_mesh_create(int nnodes, const double *xyz, int nelem, int nodes_per_elem, int dimension,
                          const int *conn, const int *celltype);
void cep_mesh_destroy(cep_mesh *m);
int cep_mesh_info(const cep_mesh *m, int *nnodes, int *nelem, int *nodes_per_elem, int *dimension,
                  int *has_celltype);
int cep_mesh_get_nodes(const cep_mesh *m, double *xyz);          /* nnodes x 3 */
int cep_mesh_get_elements(const cep_mesh *m, int *conn);         /* nelem x nodes_per_elem, 0-based */
int cep_mesh_get_celltype(const cep_mesh *m, int *celltype);     /* nnodes; CEP_E_STATE if absent */
int cep_mesh_get_targets(const cep_mesh *m, int *targets);       /* 3, 0-based */

/* Pacing-site file: *n node numbers (0-based), the first max_n copied. */
int cep_read_pacing_file(const char *path, int *nodes, int max_n, int *n);

/* Tissue on a mesh of the element types the reference's FEM assembles (Cfem::set_gpt /
   set_Phi_dPhiS, fem.cpp:830-860): S4 quadrilaterals ("square"), C3D8I hexahedra ("hex")
   and C3D4 tetrahedra ("tet"), each through Cfem::set_amat_bmat with the isoparametric map
   of every element (fem.cpp:628-661).  The same calls as a grid tissue (cep_tissue.h), with
   the conductivity given per element of the mesh (3-D meshes: cep_tissue_set_conductivity_tensor)
   and the node numbering of the mesh.  The diffusion matrices are stored in CSR.  Returns
   null for another element type or for an element whose Jacobian determinant is not
   positive at every Gauss point (inverted or folded; node order as in the reference). */
cep_tissue *cep_tissue_create_mesh(const cep_mesh *m, int model, int precision);

/* The reference's checkpoint file of every node's states, written and read back. */
int cep_tissue_write_checkpoint(const cep_tissue *t, const char *path);
int cep_tissue_read_checkpoint(cep_tissue *t, const char *path);

#ifdef __cplusplus
}
#endif
#endif
