// Explicit instantiation of the radial kernels for nz in [29, 30].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<29>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<29>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<30>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<30>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
