// Factored-DFT line passes: explicit instances of the constraint pair (x- and y-pass in one launch).
#include "hqt_spec1_core.h"

namespace hqt {
template hipError_t spec_pair_launch<double>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int, int);
template hipError_t spec_pair_launch<float>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int, int);
}  // namespace hqt
