// Explicit instantiation of the radial kernels for nz in [31, 32].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<31>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<31>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<32>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<32>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
