// March sweep launchers for FP32 slabs of 4 GB and more (per-volume buffer
// resources), every medium kind.
#include "march_launch.h"

namespace hfdtd {

HF_MARCH_RW_BIG();

}  // namespace hfdtd
