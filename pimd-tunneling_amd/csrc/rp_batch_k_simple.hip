// Batched-engine kernels for the elementwise potentials (quartic, harmonic, coupled 2D).
#include "rp_batch_kernel.h"

namespace rp {
int launch_batch_simple(const rp_pes* pes, const BatchArgs& a, int thermostat, int npoly,
                        hipStream_t s) {
  const int D = pes->ndof;
  switch (pes->kind) {
    case PES_QUARTIC:
      if (D <= 16) return launch_batch_kind<PES_QUARTIC, 16, 1>(a, thermostat, npoly, s);
      if (D <= 32) return launch_batch_kind<PES_QUARTIC, 32, 2>(a, thermostat, npoly, s);
      break;
    case PES_HARMONIC:
      if (D <= 16) return launch_batch_kind<PES_HARMONIC, 16, 1>(a, thermostat, npoly, s);
      break;
    case PES_COUPLED2D:
      return launch_batch_kind<PES_COUPLED2D, 2, 1>(a, thermostat, npoly, s);
  }
  return set_error(RP_ERR_ARG, "batched engine: potential kind %d with ndof %d not supported",
                   pes->kind, D);
}
}  // namespace rp
