// Explicit instantiation of the radial kernels for nz in [4, 12].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<4>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<4>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<5>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<5>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<6>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<6>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<7>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<7>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<8>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<8>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<9>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<9>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<10>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<10>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<11>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<11>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<12>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<12>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
