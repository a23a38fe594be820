// io.hpp -- host-side readers and writers of the reference drivers' files (C ABI in
// include/cep_io.h).  Formats follow the reference's fscanf/fprintf calls cited there.
#pragma once
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <string>

#include "../../include/cep_io.h"

extern "C" {

int cep_read_ic50_csv(const char *path, double *rows, int max_rows, int *nrows)
{
  if (!path || !rows || max_rows < 0 || !nrows) return CEP_E_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return CEP_E_ARG;
  char line[1024];
  *nrows = 0;
  if (!fgets(line, sizeof(line), f)) { fclose(f); return 0; }   // header
  while (*nrows < max_rows && fgets(line, sizeof(line), f)) {
    double *r = rows + (size_t)*nrows * 14;
    for (int k = 0; k < 14; ++k) r[k] = 0.0;
    int k = 0;
    char *save = nullptr;
    for (char *tok = strtok_r(line, ",", &save); tok && k < 14; tok = strtok_r(nullptr, ",", &save))
      r[k++] = strtod(tok, nullptr);
    if (k == 0) continue;
    ++*nrows;
  }
  fclose(f);
  return 0;
}

int cep_read_states_file(const char *path, double *states, int max_n, int *n)
{
  if (!path || !states || !n) return CEP_E_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return CEP_E_ARG;
  char line[256];
  *n = 0;
  while (*n < max_n && fgets(line, sizeof(line), f)) states[(*n)++] = strtod(line, nullptr);
  fclose(f);
  return 0;
}

int cep_write_states_file(const char *path, const double *states, int n)
{
  if (!path || !states || n < 0) return CEP_E_ARG;
  FILE *f = fopen(path, "w");
  if (!f) return CEP_E_ARG;
  for (int i = 0; i < n; ++i) fprintf(f, "%lf\n", states[i]);
  fclose(f);
  return 0;
}

void cep_param_defaults(cep_param_t *p)
{
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->simulation_mode = 0;
  p->is_dutta = 1;
  p->is_print_graph = 1;
  p->is_using_output = 0;
  p->bcl = 2000.;
  p->pace_max = 1000;
  p->celltype = 0;
  p->last_drug_check_pace = 250;
  p->dt = 0.1;
  p->dt_write = 2.0;
  p->inet_vm_threshold = -88.0;
  snprintf(p->drug_name, sizeof(p->drug_name), "%s", "bepridil");
  snprintf(p->concs, sizeof(p->concs), "%s", "99.0");
}

int cep_read_input_deck(const char *path, cep_param_t *p)
{
  if (!path || !p) return CEP_E_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return CEP_E_ARG;
  char line[256], key[100], value[100];
  while (fgets(line, sizeof(line), f)) {
    // "Key = value": the key, a separator word, the value
    if (sscanf(line, "%99s %*s %99s", key, value) != 2) continue;
    if (!strcasecmp(key, "Simulation_Mode")) p->simulation_mode = (int)strtod(value, nullptr);
    else if (!strcasecmp(key, "Celltype")) p->celltype = (int)strtod(value, nullptr);
    else if (!strcasecmp(key, "Is_Dutta")) p->is_dutta = (int)strtol(value, nullptr, 10);
    else if (!strcasecmp(key, "Is_Print_Graph")) p->is_print_graph = (int)strtol(value, nullptr, 10);
    else if (!strcasecmp(key, "Is_Using_Output")) p->is_using_output = (int)strtol(value, nullptr, 10);
    else if (!strcasecmp(key, "Basic_Cycle_Length")) p->bcl = strtod(value, nullptr);
    else if (!strcasecmp(key, "Number_of_Pacing")) p->pace_max = (int)strtol(value, nullptr, 10);
    else if (!strcasecmp(key, "Last_Drug_Check_Pace")) p->last_drug_check_pace = (int)strtol(value, nullptr, 10);
    else if (!strcasecmp(key, "Time_Step")) p->dt = strtod(value, nullptr);
    else if (!strcasecmp(key, "Write_Step")) p->dt_write = strtod(value, nullptr);
    else if (!strcasecmp(key, "Drug_Name")) snprintf(p->drug_name, sizeof(p->drug_name), "%s", value);
    else if (!strcasecmp(key, "Inet_Vm_Threshold")) p->inet_vm_threshold = strtod(value, nullptr);
    else if (!strcasecmp(key, "Concentrations")) snprintf(p->concs, sizeof(p->concs), "%s", value);
  }
  fclose(f);
  return 0;
}

int cep_param_concentrations(const cep_param_t *p, double *concs, int max_n, int *n)
{
  if (!p || !concs || !n) return CEP_E_ARG;
  char buf[sizeof(p->concs)];
  memcpy(buf, p->concs, sizeof(buf));
  buf[sizeof(buf) - 1] = 0;
  *n = 0;
  char *save = nullptr;
  for (char *tok = strtok_r(buf, ",", &save); tok && *n < max_n; tok = strtok_r(nullptr, ",", &save))
    concs[(*n)++] = strtod(tok, nullptr);
  return 0;
}

static FILE *open_append(const char *prefix, const char *suffix)
{
  const std::string path = std::string(prefix) + suffix;
  return fopen(path.c_str(), "a");
}

int cep_write_cipa_features(const char *prefix, const int *sample_ids, const double *bm, const int *status, int n,
                            const double *inal_ctl, const double *ical_ctl, int write_header)
{
  if (!prefix || !sample_ids || !bm || n < 0) return CEP_E_ARG;
  FILE *fa = open_append(prefix, "_ap_profile.plt"), *fc = open_append(prefix, "_ca_profile.plt");
  FILE *fq = open_append(prefix, "_qni.plt");
  if (!fa || !fc || !fq) {
    if (fa) fclose(fa);
    if (fc) fclose(fc);
    if (fq) fclose(fq);
    return CEP_E_ARG;
  }
  if (write_header) {
    fprintf(fa, "%s %s %s %s %s %s %s %s %s\n", "Sample_ID", "Dvm/Dt_Repol", "Max_Dvm/Dt", "Vm_Peak", "Vm_Resting",
            "APD90", "APD50", "APDTri", "Steepest_Pace");
    fprintf(fc, "%s %s %s %s %s %s %s\n", "Sample_ID", "Ca_Peak", "Ca_Diastole", "CaD90", "CaD50", "Catri",
            "Steepest_Pace");
    fprintf(fq, "%s %s %s\n", "Sample_ID", "Qnet", "Qinward");
  }
  for (int i = 0; i < n; ++i) {
    const double *r = bm + (size_t)i * CEP_NBM;
    const int id = sample_ids[i];
    if (status && status[i] != 0) {
      fprintf(fa, "%d CVODE_ERR\n", id);
      fprintf(fc, "%d CVODE_ERR\n", id);
      fprintf(fq, "%d CVODE_ERR\n", id);
      continue;
    }
    const int steep = (int)r[CEP_BM_STEEPEST_PACE];
    fprintf(fa, "%d %lf %lf %lf %lf %lf %lf %lf %d\n", id, r[CEP_BM_DVMDT_REPOL], r[CEP_BM_DVMDT_MAX],
            r[CEP_BM_VM_PEAK], r[CEP_BM_VM_DIA], r[CEP_BM_APD90], r[CEP_BM_APD50],
            r[CEP_BM_APD90] - r[CEP_BM_APD50], steep);
    fprintf(fc, "%d %lf %lf %lf %lf %lf %d\n", id, r[CEP_BM_CA_PEAK], r[CEP_BM_CA_DIA], r[CEP_BM_CAD90],
            r[CEP_BM_CAD50], r[CEP_BM_CAD90] - r[CEP_BM_CAD50], steep);
    const double qin = (inal_ctl && ical_ctl)
                           ? ((r[CEP_BM_INAL_AUC] / inal_ctl[i]) + (r[CEP_BM_ICAL_AUC] / ical_ctl[i])) * 0.5
                           : 0.0;
    fprintf(fq, "%d %lf %lf\n", id, r[CEP_BM_QNET] / 1000.0, qin);
  }
  fclose(fa);
  fclose(fc);
  fclose(fq);
  return 0;
}

int cep_write_ord11_qni(const char *prefix, const int *sample_ids, const double *bm11, int n, const double *ctl,
                        int write_header)
{
  if (!prefix || !sample_ids || !bm11 || n < 0) return CEP_E_ARG;
  FILE *fq = open_append(prefix, "_qni.plt");
  if (!fq) return CEP_E_ARG;
  if (write_header)
    fprintf(fq, "%s,%s,%s,%s,%s,%s,%s\n", "Sample_ID", "Qnet_AP", "Qnet4_AP", "Qinward_CL", "Qnet_CL", "Qnet4_CL",
            "Qinward_CL");
  for (int i = 0; i < n; ++i) {
    const double *r = bm11 + (size_t)i * CEP_NBM;
    double qin = 0.0;
    if (ctl) {
      const double *c = ctl + (size_t)i * CEP_NBM;
      qin = ((r[CEP_BM11_INAL_AUC_CL] / c[CEP_BM11_INAL_AUC_CL]) + (r[CEP_BM11_ICAL_AUC_CL] / c[CEP_BM11_ICAL_AUC_CL])) * 0.5;
    }
    fprintf(fq, "%hu,%lf,%lf,%lf,%lf,%lf,%lf\n", (unsigned short)sample_ids[i], r[CEP_BM11_QNET_AP],
            r[CEP_BM11_QNET4_AP], qin, r[CEP_BM11_QNET_CL], r[CEP_BM11_QNET4_CL], qin);
  }
  fclose(fq);
  return 0;
}

int cep_write_time_series(const char *prefix, int sample_id, const double *t, const double *vm, const double *cai,
                          const double *ires, int n, const int *rep_pace, const double *rep_val, int nrep,
                          int write_header)
{
  if (!prefix || n < 0 || nrep < 0 || (n > 0 && (!t || !vm || !cai || !ires)) || (nrep > 0 && (!rep_pace || !rep_val)))
    return CEP_E_ARG;
  const std::string base(prefix), id = std::to_string(sample_id);
  FILE *fv = fopen((base + "_vmcheck_smp" + id + ".plt").c_str(), "w");
  FILE *fc = fopen((base + "_ca_i_smp" + id + ".plt").c_str(), "w");
  FILE *fi = fopen((base + "_ires_smp" + id + ".plt").c_str(), "w");
  FILE *fr = fopen((base + "_max_dvdt_repol_smp" + id + ".plt").c_str(), "w");
  if (!fv || !fc || !fi || !fr) {
    for (FILE *f : {fv, fc, fi, fr}) if (f) fclose(f);
    return CEP_E_ARG;
  }
  if (write_header) {
    fprintf(fv, "%s %s\n", "TIME", "Vm");
    fprintf(fc, "%s %s\n", "TIME", "cai");
    fprintf(fi, "%s %s %s %s %s %s %s\n", "TIME", "INaL", "ICaL", "Ito", "IKr", "IKs", "IK1");
  }
  for (int k = 0; k < nrep; ++k) fprintf(fr, "%d %lf\n", rep_pace[k], rep_val[k]);
  for (int k = 0; k < n; ++k) fprintf(fc, "%lf %lf\n", t[k], cai[k]);
  for (int k = 0; k < n; ++k) fprintf(fv, "%lf %lf\n", t[k], vm[k]);
  for (int k = 0; k < n; ++k) {
    const double *r = ires + (size_t)k * 6;
    fprintf(fi, "%lf %lf %lf %lf %lf %lf %lf\n", t[k], r[0], r[1], r[2], r[3], r[4], r[5]);
  }
  for (FILE *f : {fv, fc, fi, fr}) fclose(f);
  return 0;
}

}  // extern "C"
