// Batched-engine kernels for the EVB double well.
#include "rp_batch_kernel.h"

namespace rp {
int launch_batch_evb(const rp_pes* pes, const BatchArgs& a, int thermostat, int npoly,
                     hipStream_t s) {
  const int D = pes->ndof;
  if (D <= 16) return launch_batch_kind<PES_EVB, 16, 1>(a, thermostat, npoly, s);
  if (D == 27) return launch_batch_kind<PES_EVB, 27, 2>(a, thermostat, npoly, s);
  if (D <= 32) return launch_batch_kind<PES_EVB, 32, 2>(a, thermostat, npoly, s);
  return set_error(RP_ERR_ARG, "batched engine: EVB with ndof %d > 32 not supported", D);
}
}  // namespace rp
