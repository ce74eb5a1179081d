// Batched march sweep launchers: FP32, one row per wavefront, 16-column
// tiles.
#include "batch_march.h"

namespace hfdtd {

HF_BATCH_MARCH_F32(, 1, 16);

}  // namespace hfdtd
