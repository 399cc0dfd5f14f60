// Instantiates the batched Cholesky kernels for padded size 32.
#include "chol_kernels.hpp"

NROCM_CHOL_DEFINE(32)
