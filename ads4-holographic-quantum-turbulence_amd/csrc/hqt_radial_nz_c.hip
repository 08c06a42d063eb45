// Explicit instantiation of the radial kernels for nz in [18, 20].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<18>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<18>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<19>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<19>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<20>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<20>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
