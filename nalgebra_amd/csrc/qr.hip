// Householder QR in nalgebra's representation and batched tall least squares.
#include "internal_launch.hpp"

extern "C" {

int nrocm_dgeqrf(nrocm_handle h, int, int, double *, int, double *)
{
    return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "dgeqrf: not built yet");
}
int nrocm_sgeqrf(nrocm_handle h, int, int, float *, int, float *)
{
    return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "sgeqrf: not built yet");
}
int nrocm_dqr_qtmul(nrocm_handle h, int, int, const double *, int, const double *, int, double *, int)
{
    return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "dqr_qtmul: not built yet");
}
int nrocm_dqr_rsolve(nrocm_handle h, int, const double *, int, const double *, int, double *, int, int *)
{
    return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "dqr_rsolve: not built yet");
}
int nrocm_sgels_batched(nrocm_handle h, nrocm_batch, int, float *const *, float *const *, double *, int, int *)
{
    return nrocm_fail(h, NROCM_ERR_UNSUPPORTED, "sgels_batched: not built yet");
}

}  // extern "C"
