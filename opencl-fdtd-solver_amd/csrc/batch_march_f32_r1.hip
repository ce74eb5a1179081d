// Batched march sweep launchers: FP32, one row per wavefront, 64- and
// 32-column tiles.
#include "batch_march.h"

namespace hfdtd {

HF_BATCH_MARCH_F32(, 1, 64);
HF_BATCH_MARCH_F32(, 1, 32);

}  // namespace hfdtd
