/*
 * cep_io.h -- the reference drivers' input and output files (libcepgpu.so, host code).
 *
 * Lets a driver built on the batch ABI read the reference's inputs and write its result files
 * in the same formats:
 *   IC50 sample CSV        DrugSimulation/main.cpp:308-336 (get_IC50_data_from_file)
 *   EDISON input deck      DrugSimulation/modules/glob_funct.cpp:60-133, defaults param.cpp:6-24
 *   state dump             output_orudy.dat, read CEPCell_NEO/modules/cipa_drug_bench.cpp:102-113,
 *                          written cipa_drug_full_bench.cpp:353 ("%lf" per line)
 *   CiPA feature files     CEPCell_NEO/modules/cipa_drug_bench.cpp:177-199,354-381
 *   ORd 2011 qNet file     DrugSimulation/modules/drug_sim.cpp:135-139,430-447
 * All functions return 0 or a negative CEP_E_* code; none touches the GPU.
 */
#ifndef CEP_IO_H
#define CEP_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* Rows of the IC50 sample CSV after its header line: up to 14 comma-separated values per
   row, each read with strtod, so "NA" (and anything else non-numeric) is 0 = no block.
   rows: max_rows x 14 (missing trailing values are 0); *nrows = rows read. */
int cep_read_ic50_csv(const char *path, double *rows, int max_rows, int *nrows);

/* One value per line (the steady-state dump).  *n = values read (at most max_n). */
int cep_read_states_file(const char *path, double *states, int max_n, int *n);
int cep_write_states_file(const char *path, const double *states, int n);

/* The driver parameters of the EDISON deck ("Key = value" lines, keys case-insensitive) */
typedef struct {
  int simulation_mode, celltype, is_dutta, is_print_graph, is_using_output;
  int pace_max, last_drug_check_pace;
  double bcl, dt, dt_write, inet_vm_threshold;
  char drug_name[100];
  char concs[100];          /* comma-separated concentrations, as in the deck */
} cep_param_t;
/* the reference's param_t::init values */
void cep_param_defaults(cep_param_t *p);
/* overwrite the fields named in the deck; unknown keys are ignored */
int cep_read_input_deck(const char *path, cep_param_t *p);
/* the deck's concentration list as numbers; *n = values (at most max_n) */
int cep_param_concentrations(const cep_param_t *p, double *concs, int max_n, int *n);

/* Append one concentration's CiPA features (records of cep_get_biomarkers, n x CEP_NBM) to
   <prefix>_ap_profile.plt, <prefix>_ca_profile.plt and <prefix>_qni.plt, with the reference's
   header lines when write_header != 0.  qNet is written /1000; qInward uses the control
   AUCs inal_ctl / ical_ctl (n each), or 0 when they are null (the control run).  status
   (n, may be null): a nonzero status writes the reference's "<id> CVODE_ERR" line. */
int cep_write_cipa_features(const char *prefix, const int *sample_ids, const double *bm, const int *status,
                            int n, const double *inal_ctl, const double *ical_ctl, int write_header);

/* Append ORd 2011 qNet rows (records CEP_BM11_*, n x CEP_NBM) to <prefix>_qni.plt:
   id, qnet_ap, qnet4_ap, qinward_cl, qnet_cl, qnet4_cl, qinward_cl, with qinward from the
   control records ctl (n x CEP_NBM, null for the control run). */
int cep_write_ord11_qni(const char *prefix, const int *sample_ids, const double *bm11, int n,
                        const double *ctl, int write_header);

/* Write one sample's time series in the reference's Is_Print_Graph files (opened "w"):
   <prefix>_vmcheck_smp<id>.plt    "TIME Vm"
   <prefix>_ca_i_smp<id>.plt       "TIME cai"
   <prefix>_ires_smp<id>.plt       "TIME INaL ICaL Ito IKr IKs IK1"
   <prefix>_max_dvdt_repol_smp<id>.plt  "<pace> <dvmdt_repol>" per line, no header
   (CEPCell_NEO/modules/cipa_drug_bench.cpp:168-177,189-193,262,338-350; "%lf" fields; the
   headers only when write_header != 0, as the reference writes them for group 0).
   ires: n x 6; rep_pace / rep_val: nrep entries (may be null with nrep = 0). */
int cep_write_time_series(const char *prefix, int sample_id, const double *t, const double *vm, const double *cai,
                          const double *ires, int n, const int *rep_pace, const double *rep_val, int nrep,
                          int write_header);

#ifdef __cplusplus
}
#endif
#endif
