// Instantiates the batched Cholesky kernels for padded size 64.
#include "chol_kernels.hpp"
#include "chol_tile.hpp"

NROCM_CHOL_DEFINE(64)
NROCM_CHOL_TILE_DEFINE(64)
