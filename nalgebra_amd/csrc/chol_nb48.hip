// Instantiates the batched Cholesky kernels for padded size 48.
#include "chol_kernels.hpp"
#include "chol_tile.hpp"

NROCM_CHOL_DEFINE(48)
NROCM_CHOL_TILE_DEFINE(48)
