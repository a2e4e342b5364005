// Batch backward compositor instances for the 16x16 binning tile (every
// backward window depth and sync threshold for n_samp = 1, the defaults for
// n_samp 2 and 3).  The kernels are in batch_common.h.
#include "batch_common.h"

namespace svr {

SVR_DEFINE_BATCH_BWD_RUN(Tile16)

}  // namespace svr
