// Batched march sweep launchers: FP32, two rows per wavefront.
#include "batch_march.h"

namespace hfdtd {

HF_BATCH_MARCH_F32(, 2, 64);

}  // namespace hfdtd
