// Explicit instantiation of the radial kernels for nz in [13, 17].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<13>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<13>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<14>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<14>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<15>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<15>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<16>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<16>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<17>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<17>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
