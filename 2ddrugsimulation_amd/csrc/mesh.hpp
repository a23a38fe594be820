// mesh.hpp -- the reference tissue driver's mesh and pacing-site files (C ABI in
// include/cep_mesh.h; host code).  The reader follows the section order of Cmesh::loadInp
// (mesh.cpp:59-183) and the pacing reader Cheart::loadPacingsites (heart.cpp:966-998).
#pragma once
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/cep_mesh.h"

struct cep_mesh {
  int nn = 0, ne = 0, npe = 0, dim = 0;
  int targets[3] = { 0, 0, 0 };
  std::vector<double> xyz;        // nn x 3
  std::vector<int> conn;          // ne x npe, 0-based
  std::vector<int> celltype;      // nn or empty
};

namespace cep {
namespace meshio {

// element type keyword -> nodes per element and dimension (mesh.cpp:93-122)
static bool element_type(const char *tok, int &npe, int &dim)
{
  while (*tok == ' ' || *tok == '\t') ++tok;
  struct T { const char *key; int npe, dim; };
  // longer keys first: "TYPE=S4" would otherwise not be told from "TYPE=S4R"-like prefixes
  static const T types[] = { { "TYPE=C3D8I", 8, 3 }, { "TYPE=C3D4", 4, 3 }, { "TYPE=T3D2", 2, 3 },
                             { "TYPE=S2", 2, 1 },    { "TYPE=S4", 4, 2 },   { "TYPE=S3", 3, 3 } };
  for (const T &t : types)
    if (!strncmp(tok, t.key, strlen(t.key))) { npe = t.npe; dim = t.dim; return true; }
  return false;
}

static bool next_line(FILE *f, char *buf, int n) { return fgets(buf, n, f) != nullptr; }

// comma-separated fields of a line (the leading id first)
static int split_fields(char *line, char **out, int max)
{
  int k = 0;
  char *save = nullptr;
  for (char *tok = strtok_r(line, ",", &save); tok && k < max; tok = strtok_r(nullptr, ",", &save)) out[k++] = tok;
  return k;
}

static cep_mesh *read_inp(const char *path)
{
  FILE *f = fopen(path, "r");
  if (!f) return nullptr;
  char line[1024];
  cep_mesh *m = new cep_mesh();
  bool ok = false;
  do {
    // header up to the "CML" line, then the counts and the three target nodes
    bool cml = false;
    while (next_line(f, line, sizeof(line)))
      if (!strncmp(line, "CML", 3)) { cml = true; break; }
    if (!cml || !next_line(f, line, sizeof(line))) break;
    int t0, t1, t2;
    if (sscanf(line, "%d %d %d %d %d", &m->nn, &m->ne, &t0, &t1, &t2) != 5 || m->nn <= 0 || m->ne <= 0) break;
    m->targets[0] = t0 - 1; m->targets[1] = t1 - 1; m->targets[2] = t2 - 1;
    bool node = false;
    while (next_line(f, line, sizeof(line)))
      if (!strncmp(line, "*NODE", 5)) { node = true; break; }
    if (!node) break;
    m->xyz.assign((size_t)m->nn * 3, 0.0);
    bool bad = false;
    for (int i = 0; i < m->nn && !bad; ++i) {
      char *fld[8];
      if (!next_line(f, line, sizeof(line)) || split_fields(line, fld, 8) < 4) { bad = true; break; }
      for (int j = 0; j < 3; ++j) m->xyz[(size_t)i * 3 + j] = atof(fld[1 + j]);
    }
    if (bad) break;
    // "**HWCOLOR ..." then "*ELEMENT, TYPE=..."
    if (!next_line(f, line, sizeof(line)) || !next_line(f, line, sizeof(line))) break;
    char *fld[16];
    if (strncmp(line, "*ELEMENT", 8) || split_fields(line, fld, 16) < 2 || !element_type(fld[1], m->npe, m->dim)) break;
    m->conn.assign((size_t)m->ne * m->npe, 0);
    for (int e = 0; e < m->ne && !bad; ++e) {
      if (!next_line(f, line, sizeof(line)) || split_fields(line, fld, 16) < m->npe + 1) { bad = true; break; }
      for (int j = 0; j < m->npe; ++j) {
        const int v = atoi(fld[1 + j]) - 1;
        if (v < 0 || v >= m->nn) { bad = true; break; }
        m->conn[(size_t)e * m->npe + j] = v;
      }
    }
    if (bad || !next_line(f, line, sizeof(line))) break;
    if (!strncmp(line, "CELLTYPE", 8)) {
      m->celltype.assign(m->nn, 0);
      for (int i = 0; i < m->nn && !bad; ++i) {
        if (!next_line(f, line, sizeof(line))) { bad = true; break; }
        m->celltype[i] = atoi(line);
      }
      if (bad || !next_line(f, line, sizeof(line))) break;
    }
    ok = !strncmp(line, "eof", 3);
  } while (false);
  fclose(f);
  if (!ok) { delete m; return nullptr; }
  return m;
}

}  // namespace meshio
}  // namespace cep

extern "C" {

cep_mesh *cep_mesh_read_inp(const char *path)
{
  if (!path) return nullptr;
  return cep::meshio::read_inp(path);
}

cep_mesh *cep_mesh_create(int nnodes, const double *xyz, int nelem, int nodes_per_elem, int dimension,
                          const int *conn, const int *celltype)
{
  if (nnodes <= 0 || nelem <= 0 || !xyz || !conn || nodes_per_elem < 2 || nodes_per_elem > 8 || dimension < 1 ||
      dimension > 3)
    return nullptr;
  cep_mesh *m = new cep_mesh();
  m->nn = nnodes; m->ne = nelem; m->npe = nodes_per_elem; m->dim = dimension;
  m->xyz.assign(xyz, xyz + (size_t)nnodes * 3);
  m->conn.assign(conn, conn + (size_t)nelem * nodes_per_elem);
  for (int v : m->conn)
    if (v < 0 || v >= nnodes) { delete m; return nullptr; }
  if (celltype) m->celltype.assign(celltype, celltype + nnodes);
  return m;
}

void cep_mesh_destroy(cep_mesh *m) { delete m; }

int cep_mesh_info(const cep_mesh *m, int *nnodes, int *nelem, int *nodes_per_elem, int *dimension, int *has_celltype)
{
  if (!m) return CEP_E_ARG;
  if (nnodes) *nnodes = m->nn;
  if (nelem) *nelem = m->ne;
  if (nodes_per_elem) *nodes_per_elem = m->npe;
  if (dimension) *dimension = m->dim;
  if (has_celltype) *has_celltype = m->celltype.empty() ? 0 : 1;
  return 0;
}

int cep_mesh_get_nodes(const cep_mesh *m, double *xyz)
{
  if (!m || !xyz) return CEP_E_ARG;
  memcpy(xyz, m->xyz.data(), m->xyz.size() * sizeof(double));
  return 0;
}

int cep_mesh_get_elements(const cep_mesh *m, int *conn)
{
  if (!m || !conn) return CEP_E_ARG;
  memcpy(conn, m->conn.data(), m->conn.size() * sizeof(int));
  return 0;
}

int cep_mesh_get_celltype(const cep_mesh *m, int *celltype)
{
  if (!m || !celltype) return CEP_E_ARG;
  if (m->celltype.empty()) return CEP_E_STATE;
  memcpy(celltype, m->celltype.data(), m->celltype.size() * sizeof(int));
  return 0;
}

int cep_mesh_get_targets(const cep_mesh *m, int *targets)
{
  if (!m || !targets) return CEP_E_ARG;
  for (int k = 0; k < 3; ++k) targets[k] = m->targets[k];
  return 0;
}

int cep_read_pacing_file(const char *path, int *nodes, int max_n, int *n)
{
  if (!path || !n || max_n < 0 || (max_n > 0 && !nodes)) return CEP_E_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return CEP_E_ARG;
  int cnt = 0;
  if (fscanf(f, "%d", &cnt) != 1 || cnt < 0) { fclose(f); return CEP_E_ARG; }
  *n = cnt;
  for (int k = 0; k < cnt; ++k) {
    int v = 0;
    if (fscanf(f, "%d", &v) != 1) { fclose(f); return CEP_E_ARG; }
    if (k < max_n) nodes[k] = v - 1;    // node 1 of the file is node 0 here
  }
  fclose(f);
  return 0;
}

}  // extern "C"
