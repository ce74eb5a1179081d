// March sweep launchers: FP64 (one row per wavefront), every medium kind.
#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_SET(, double, 1, 64);

}  // namespace hfdtd
