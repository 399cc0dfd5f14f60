// Mixed-precision LU + FP64 iterative refinement.
#include "internal_launch.hpp"

extern "C" {

int nrocm_dgesv_ir(nrocm_handle h, int, const double *, int, const double *, double *, int, double, int, int *,
                   double *)
{
    return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "dgesv_ir: not built yet");
}

}  // extern "C"
