// Explicit instantiation of the radial kernels for nz in [21, 23].
#include "hqt_radial_core.h"

namespace hqt {
template hipError_t zstage_launch_nz<21>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<21>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<22>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<22>(const ZPost*, Members, int, int, hipStream_t);
template hipError_t zstage_launch_nz<23>(const ZStage*, Members, int, int, int, hipStream_t);
template hipError_t zpost_launch_nz<23>(const ZPost*, Members, int, int, hipStream_t);
}  // namespace hqt
