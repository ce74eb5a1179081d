// March sweep launchers: FP32, tiles 32 columns wide, every medium kind.
#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_SET(, float, 1, 32);
HF_MARCH_RW_MAT(, float, 1, 32);

}  // namespace hfdtd
