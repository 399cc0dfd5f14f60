// Instantiates the batched Cholesky kernels for padded size 24.
#include "chol_kernels.hpp"
#include "chol_tile.hpp"
#include "chol_tile2.hpp"

NROCM_CHOL_DEFINE(24)
NROCM_CHOL_TILE_DEFINE(24)
NROCM_CHOL_TILE2_DEFINE(24)
