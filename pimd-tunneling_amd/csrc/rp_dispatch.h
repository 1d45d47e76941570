// rp_dispatch.h -- compile-time selection of a potential's code path.
//
// Kernels that evaluate a potential are templated on <KIND, MAXD>: KIND picks the
// functional form, MAXD the register-array bound (ndof <= MAXD).  For WATER,
// MAXD = 9 * molecules.
#pragma once
#include "rp_internal.h"

namespace rp {

template <int KIND, int MAXD, class S>
__device__ __forceinline__ S eval_kind(const PesDev& P, const S* x, S* g) {
  S v;
  if constexpr (KIND == PES_QUARTIC)
    v = quartic_eval<MAXD>(P, x, g);
  else if constexpr (KIND == PES_HARMONIC)
    v = harmonic_eval<MAXD>(P, x, g);
  else if constexpr (KIND == PES_COUPLED2D)
    v = coupled_eval(P, x, g);
  else if constexpr (KIND == PES_EVB)
    v = evb_eval<MAXD>(P, x, g);
  else
    v = water_eval<MAXD / 9>(P, x, g);
  return v - P.v0;
}

// Calls f.template operator()<KIND, MAXD>() for the potential; RP_ERR_ARG if unsupported.
template <class F>
int dispatch_pes(const rp_pes* pes, F&& f) {
  const int D = pes->ndof;
  switch (pes->kind) {
    case PES_QUARTIC:
      if (D <= 16) return f.template operator()<PES_QUARTIC, 16>();
      if (D <= 32) return f.template operator()<PES_QUARTIC, 32>();
      if (D <= 64) return f.template operator()<PES_QUARTIC, 64>();
      break;
    case PES_HARMONIC:
      if (D <= 16) return f.template operator()<PES_HARMONIC, 16>();
      if (D <= 32) return f.template operator()<PES_HARMONIC, 32>();
      if (D <= 64) return f.template operator()<PES_HARMONIC, 64>();
      break;
    case PES_COUPLED2D:
      if (D == 2) return f.template operator()<PES_COUPLED2D, 2>();
      break;
    case PES_EVB:
      if (D <= 16) return f.template operator()<PES_EVB, 16>();
      if (D == 27) return f.template operator()<PES_EVB, 27>();
      if (D <= 32) return f.template operator()<PES_EVB, 32>();
      break;
    case PES_HOST:
      return set_error(RP_ERR_ARG, "a potential that lives on the host has no device kernel "
                                   "on this path (the batched engine and the parameter "
                                   "sensitivities need a device potential)");
    case PES_WATER:
      if (pes->ndim == 3 && pes->natom == 3) return f.template operator()<PES_WATER, 9>();
      if (pes->ndim == 3 && pes->natom == 6) return f.template operator()<PES_WATER, 18>();
      if (pes->ndim == 3 && pes->natom == 18) return f.template operator()<PES_WATER, 54>();
      break;
  }
  return set_error(RP_ERR_ARG, "potential kind %d with ndim=%d natom=%d is not supported",
                   pes->kind, pes->ndim, pes->natom);
}

}  // namespace rp
