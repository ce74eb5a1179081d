// March sweep launchers: FP32, tiles 16 columns wide, every medium kind.
#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_SET(, float, 1, 16);
HF_MARCH_RW_MAT(, float, 1, 16);

}  // namespace hfdtd
