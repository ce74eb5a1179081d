// Batched march sweep launchers: FP64, one row per wavefront.
#include "batch_march.h"

namespace hfdtd {

HF_BATCH_MARCH_F64(, 64);

}  // namespace hfdtd
