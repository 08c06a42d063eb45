// Folded line passes: explicit instances of the float x-direction family.
#include "hqt_fold_core.h"

namespace hqt {
template hipError_t fold_launch_td<float, 0>(const LPass*, Members, dim3, int, int, int, size_t, hipStream_t, int, int);
}  // namespace hqt
