// population_adapter.cpp -- the reference-side adapter of INTEGRATION.md, compiled against
// include/cep_cellsim.h and linked with libcepgpu.so by tests/test_adapter.py.
//
// run_population_on_gpu: the drug loop of CEP7Cell/modules/drugs_herg.cpp:112-303 (one CVode run
// per (sample, concentration)) as one batch.
// run_drugs_herg_tail: the same loop with the reference's own per-output-point CVode(t_out) calls
// for the last `tail` output points (cep_integrate_bdf), its drugs_herg outputs (APD at v_apd90,
// qnet, qinward against the concentration-0 run).
//
// Usage: population_adapter <in> <out> (raw little-endian float64)
//   in : n_samples, n_conc, bcl, dt, stim_amp, num_pace1, tail, then n_conc concentrations, then
//        n_samples rows of 19 herg values
//   out: per job (sample-major, conc-minor): the CEP_NBM record, the CEP_NDR driver-rule record,
//        qinward, status
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cep_cellsim.h"

// rows = every (sample, conc) the reference loop would visit, in the same order
int run_population_on_gpu(const std::vector<std::vector<double>> &herg_rows,   // 19 per row:
                          // 14 IC50/Hill values (CSV order) + Kmax, Ku, n, halfmax, Vhalf
                          const std::vector<double> &concs, double bcl, double dt, double stim_amp,
                          int num_pace1, int tail, std::vector<double> &biomarkers, std::vector<double> &rules,
                          std::vector<double> &qinward, std::vector<int> &status)
{
  const int n = (int)(herg_rows.size() * concs.size());
  std::vector<double> celltype(n, 0.0), conc(n), herg((size_t)n * 19), bcls(n, bcl);
  std::vector<int> control(n);
  for (size_t s = 0, k = 0; s < herg_rows.size(); ++s)
    for (size_t c = 0; c < concs.size(); ++c, ++k) {
      conc[k] = concs[c];
      std::copy(herg_rows[s].begin(), herg_rows[s].end(), herg.begin() + k * 19);
      control[k] = (int)(s * concs.size());      // the sample's first concentration is its control
    }
  cep_batch *b = cep_batch_create(CEP_MODEL_ORD_CIPA, n, CEP_PREC_MIXED);   // or CEP_PREC_FP64
  if (!b) return -1;
  int rc = cep_set_cells(b, 0, n, celltype.data(), conc.data(), herg.data(), 19, /*gscale*/ nullptr, 0,
                         bcls.data());
  // CVodeSStolerances(1e-7, 1e-7), CVodeSetMaxStep(dt); biomarkers over the last 250 paces
  if (!rc) rc = cep_set_run_config(b, dt, 1e-7, 1e-7, 5, num_pace1, num_pace1 > 250 ? 250 : num_pace1, stim_amp);
  // the drugs_herg outputs besides the CiPA record: v_apd90 = -40 (CEP7Cell/commons/helper.cpp:21)
  if (!rc) rc = cep_set_model_option(b, CEP_OPT_DRIVER_RULES, 1.0);
  if (!rc) rc = cep_set_model_option(b, CEP_OPT_V_APD90, -40.0);
  if (!rc) rc = cep_init_consts(b);
  const long long imax = (long long)((num_pace1 * bcl) / dt);
  // all but the last `tail` output points in one launch, then CVode(t_out) per output point
  if (!rc) rc = cep_run_outputs(b, (int)(imax - tail));
  for (long long iout = imax - tail + 1; !rc && iout <= imax; ++iout) rc = cep_integrate_bdf(b, iout * dt);
  if (!rc) {
    biomarkers.resize((size_t)n * CEP_NBM);
    rc = cep_get_biomarkers(b, biomarkers.data());
  }
  if (!rc) {
    rules.resize((size_t)n * CEP_NDR);
    rc = cep_get_driver_rules(b, rules.data());
  }
  if (!rc) {
    qinward.resize(n);
    rc = cep_get_qinward(b, control.data(), qinward.data());
  }
  status.assign(n, 0);
  if (!rc) rc = cep_get_status(b, status.data());
  // status[k] != CEP_ST_OK  <=>  the reference's "CVode error ..." / "CVODE_ERR" line for job k
  if (rc) fprintf(stderr, "cep: %s\n", cep_last_error(b));
  cep_batch_destroy(b);
  return rc;
}

int main(int argc, char **argv)
{
  if (argc != 3) {
    fprintf(stderr, "usage: population_adapter in out\n");
    return 2;
  }
  std::vector<double> in;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  double x;
  while (fread(&x, sizeof(double), 1, f) == 1) in.push_back(x);
  fclose(f);
  const int ns = (int)in[0], nc = (int)in[1];
  const double bcl = in[2], dt = in[3], stim = in[4];
  const int np = (int)in[5], tail = (int)in[6];
  std::vector<double> concs(in.begin() + 7, in.begin() + 7 + nc);
  std::vector<std::vector<double>> rows(ns);
  for (int s = 0; s < ns; ++s) rows[s].assign(in.begin() + 7 + nc + 19 * s, in.begin() + 7 + nc + 19 * (s + 1));
  std::vector<double> bm, dr, qin;
  std::vector<int> st;
  const int rc = run_population_on_gpu(rows, concs, bcl, dt, stim, np, tail, bm, dr, qin, st);
  if (rc) return 1;
  FILE *o = fopen(argv[2], "wb");
  const int n = ns * nc;
  for (int k = 0; k < n; ++k) {
    fwrite(&bm[(size_t)k * CEP_NBM], sizeof(double), CEP_NBM, o);
    fwrite(&dr[(size_t)k * CEP_NDR], sizeof(double), CEP_NDR, o);
    fwrite(&qin[k], sizeof(double), 1, o);
    const double sd = st[k];
    fwrite(&sd, sizeof(double), 1, o);
  }
  fclose(o);
  return 0;
}
