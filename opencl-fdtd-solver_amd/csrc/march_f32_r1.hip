// March sweep launchers: FP32, one row per wavefront, every medium kind.
#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_SET(, float, 1, 64);
HF_MARCH_RW_MAT(, float, 1, 64);

}  // namespace hfdtd
