"""nalgebra_amd -- MI355X (gfx950) implementation of nalgebra's dense hot path.

The Python layer mirrors the reference's API for this path:

    Matrix::gemm                       -> DMatrix.gemm / gemm()
    Matrix::lu / LU::solve_mut         -> DMatrix.lu() -> LU.solve_mut()
    Matrix::cholesky / solve_mut       -> DMatrix.cholesky() -> Optional[Cholesky]
    Matrix::qr / q_tr_mul / solve_mut  -> DMatrix.qr() -> QR
    (batched, mixed precision, adjoint) -> CholeskyBatch, solve_mixed_ir, backward()

All arithmetic runs in libnalgebra_rocm.so (HIP kernels behind a C ABI);
torch supplies device memory and the stream.  There is no CPU fallback.
"""
from ._lib import EXACT, FAST, PREC_BF16, PREC_BF16X3, PREC_FP32, PREC_FP64, PREC_GMRES, PREC_WINDOW, NrocmError, header_functions
from .linalg import (
    DMatrix,
    LU,
    Cholesky,
    QR,
    PermutationSequence,
    gemm,
    handle,
    solve_mixed_ir,
    last_ir_fallback,
)
from .batched import CholeskyBatch, LeastSquaresBatch, LUBatch

__all__ = [
    "EXACT", "FAST", "PREC_FP64", "PREC_FP32", "PREC_BF16", "PREC_BF16X3", "PREC_GMRES", "PREC_WINDOW", "NrocmError", "header_functions",
    "DMatrix", "LU", "Cholesky", "QR", "PermutationSequence", "gemm", "handle", "solve_mixed_ir", "last_ir_fallback",
    "CholeskyBatch", "LeastSquaresBatch", "LUBatch",
]
