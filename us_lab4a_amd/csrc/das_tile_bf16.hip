// LDS-tiled DAS kernels for BF16 analytic-signal storage.
#include "das_common.h"

namespace mista {
namespace das {
template void launch_tile_storage<MISTA_STORAGE_BF16>(int, const TileBatch&, uint32_t, uint32_t, hipStream_t);
}  // namespace das
}  // namespace mista
