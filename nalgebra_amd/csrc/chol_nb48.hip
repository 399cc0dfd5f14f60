// Instantiates the batched Cholesky kernels for padded size 48.
#include "chol_kernels.hpp"

NROCM_CHOL_DEFINE(48)
