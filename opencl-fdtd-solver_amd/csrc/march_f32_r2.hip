// March sweep launchers: FP32, two rows per wavefront, every medium kind.
#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_SET(, float, 2, 64);
HF_MARCH_RW_MAT(, float, 2, 64);

}  // namespace hfdtd
