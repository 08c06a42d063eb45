// Folded line passes: the paired constraint launch (x- and y-pass in one grid), both precisions.
#include "hqt_fold_core.h"

namespace hqt {
template hipError_t fold_pair_launch_cons<double>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int, int);
template hipError_t fold_pair_launch_cons<float>(const LPass*, const LPass*, int, int, int, size_t, hipStream_t, int, int);
}  // namespace hqt
