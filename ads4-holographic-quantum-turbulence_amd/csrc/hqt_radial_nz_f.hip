// Explicit instantiation of the radial kernels for nz in [27, 28].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<27>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<27>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<28>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<28>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
