// Batch backward compositor instances for the 32x8 binning tile (every
// backward window depth and sync threshold for n_samp = 1, the defaults for
// n_samp 2 and 3).  The kernels are in batch_common.h.
#include "batch_common.h"

namespace svr {

using Tile32x8 = TileShape<32, 8>;
SVR_DEFINE_BATCH_BWD_RUN(Tile32x8)

}  // namespace svr
