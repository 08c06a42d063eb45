// Explicit instantiation of the radial kernels for nz in [24, 26].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<24>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<24>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<25>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<25>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<26>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<26>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
