// Factored-DFT line passes, single-region form: explicit instances of the float x-direction family, filter passes.
#include "hqt_spec1_core.h"

namespace hqt {
template hipError_t spec1_launch_td<float, 0, true>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
}  // namespace hqt
