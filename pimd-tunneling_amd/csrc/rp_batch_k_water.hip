// Batched-engine kernels for the water clusters (1, 2 or 6 molecules), on the full
// (unfolded) transforms.
#include "rp_batch_kernel.h"

namespace rp {
int launch_batch_water(const rp_pes* pes, const BatchArgs& a, int thermostat, int npoly,
                       hipStream_t s) {
  switch (pes->natom) {
    case 3: return launch_batch_kind<PES_WATER, 9, 1, false>(a, thermostat, npoly, s);
    case 6: return launch_batch_kind<PES_WATER, 18, 2, false>(a, thermostat, npoly, s);
    case 18: return launch_batch_kind<PES_WATER, 54, 4, false>(a, thermostat, npoly, s);
  }
  return set_error(RP_ERR_ARG, "batched engine: water cluster with %d atoms not supported",
                   pes->natom);
}
}  // namespace rp
