// Instantiates the batched Cholesky kernels for padded size 64.
#include "chol_kernels.hpp"

NROCM_CHOL_DEFINE(64)
