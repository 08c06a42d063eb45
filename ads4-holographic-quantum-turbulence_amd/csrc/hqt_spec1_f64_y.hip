// Factored-DFT line passes, single-region form: explicit instances of the double y-direction family.
#include "hqt_spec1_core.h"

namespace hqt {
template hipError_t spec1_launch_td<double, 1, false>(const LPass*, Members, dim3, int, size_t, hipStream_t, int, int);
}  // namespace hqt
