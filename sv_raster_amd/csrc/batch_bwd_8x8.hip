// Batch backward compositor instances for the 8x8 binning tile (every
// backward window depth and sync threshold for n_samp = 1, the defaults for
// n_samp 2 and 3).  The kernels are in batch_common.h.
#include "batch_common.h"

namespace svr {

using Tile8x8 = TileShape<8, 8>;
SVR_DEFINE_BATCH_BWD_RUN(Tile8x8)

}  // namespace svr
