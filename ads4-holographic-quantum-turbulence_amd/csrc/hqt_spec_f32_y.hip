// Factored-DFT line passes: explicit instances of the float y-direction family, RK-stage passes.
#include "hqt_spec_core.h"

namespace hqt {
template hipError_t spec_launch_td<float, 1, false>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
}  // namespace hqt
