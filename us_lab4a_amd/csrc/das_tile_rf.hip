// LDS-tiled DAS kernels for real (RF) samples.
#include "das_common.h"

namespace mista {
namespace das {
template void launch_tile_storage<kStorageReal>(int, const TileBatch&, uint32_t, uint32_t, hipStream_t);
}  // namespace das
}  // namespace mista
