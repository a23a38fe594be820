// ord11_host.hpp -- host-side inputs of the ORd 2011 cells: the per-cell record the C ABI
// fills, the initial STATES and the per-cell device parameters (DrugSimulation/cellmodels/
// Ohara_Rudy_2011.cpp ___initConsts :494-679, ___applyDutta :460-467, ___applyDrugEffect
// :469-485).  Shared by the library and the host build of the Adams integrator used by tests.
#pragma once
#include <math.h>
#include "ord11_model.hpp"

namespace cep {
namespace host {

// ---- ORd 2011 (DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp)
struct O11CellIn {
  double celltype, conc, bcl;
  double ic50[14];
};

// initial STATES of ___initConsts (:494-679)
inline constexpr double kO11Init[41] = {
  -87, 0, 1e-4, 7, 7, 145, 145, 1.2, 1.2, 1e-4, 0, 1, 1, 1, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 1,
  0, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 1, 0, 0 };

// per-cell parameters: endocardial constants, the Dutta scaling of the reference (slots 96,
// 98, 100, 81 of CONSTANTS = GNaL, Aff, tjca, H), then the Hill block of ___applyDrugEffect
inline void o11_params(const O11CellIn &in, bool dutta, double stim_scale, double *p)
{
  using namespace ord11;
  auto block = [&](int k) {
    const double ic = in.ic50[2 * k], h = in.ic50[2 * k + 1];
    return (ic > 10E-14 && h > 10E-14) ? 1. / (1. + pow(in.conc / ic, h)) : 1.;
  };
  p[P_PCA] = PCa_b * block(0);
  p[P_GK1] = GK1_b * block(1);
  p[P_GKS] = GKs_b * block(2);
  p[P_GNA] = GNa * block(3);
  p[P_GNAL] = (dutta ? GNaL_b * 1.013 : GNaL_b) * block(4);
  p[P_GTO] = Gto_b * block(5);
  p[P_GKR] = GKr_b * block(6);
  p[P_AFF] = dutta ? 0.6 * 1.870 : 0.6;
  p[P_TJCA] = dutta ? 75.0 * 1.698 : 75.0;
  p[P_H] = dutta ? 1e-7 * 1.007 : 1e-7;
  p[P_BCL] = in.bcl;
  p[P_STIM_AMP] = stim_amp * stim_scale;
}

}  // namespace host
}  // namespace cep
