// Line passes: explicit instances of the float x-direction family.
#include "hqt_line_core.h"

namespace hqt {
template hipError_t line_launch_td<float, 0>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t);
}  // namespace hqt
