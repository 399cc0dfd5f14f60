// Instantiates the batched Cholesky kernels for padded size 40.
#include "chol_kernels.hpp"
#include "chol_tile.hpp"

NROCM_CHOL_DEFINE(40)
NROCM_CHOL_TILE_DEFINE(40)
