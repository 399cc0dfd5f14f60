"""Batched uneven-size solvers: Cholesky + solve with its adjoint, LU + solve,
and tall least squares.  Members are packed back to back in one device buffer each
(a Vec<DMatrix> flattened), and the C ABI receives device arrays of
per-member pointers.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from ._lib import EXACT, check, lib
from .linalg import _ptr, handle


def _pointer_array(base: torch.Tensor, offsets: np.ndarray) -> torch.Tensor:
    """Device array of element pointers base + offsets (in elements)."""
    addrs = base.data_ptr() + offsets.astype(np.int64) * base.element_size()
    return torch.from_numpy(addrs.astype(np.int64)).to(base.device)


def _pointer_list(ts: Sequence[torch.Tensor], device) -> torch.Tensor:
    return torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device=device)


class CholeskyBatch:
    """count SPD systems A_j (n_j x n_j) with B_j (n_j x nrhs), factored and
    solved in one graph-captured launch sequence (nrocm_dpotrf_potrs_batched)."""

    def __init__(self, ns: Sequence[int], nrhs: int, device="cuda"):
        self.ns = np.ascontiguousarray(ns, dtype=np.int32)
        self.count = len(self.ns)
        self.nrhs = int(nrhs)
        self.device = torch.device(device)
        n64 = self.ns.astype(np.int64)
        self.offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
        self.offB = np.concatenate([[0], np.cumsum(n64 * self.nrhs)[:-1]]).astype(np.int64)
        self.sizeA = int((n64 * n64).sum())
        self.sizeB = int((n64 * self.nrhs).sum())
        self._h = handle()
        bp = C.c_void_p()
        check("nrocm_batch_create",
              lib().nrocm_batch_create(self._h, self.count,
                                       self.ns.ctypes.data_as(C.POINTER(C.c_int)), self.nrhs,
                                       C.byref(bp)), self._h)
        self._batch = bp.value
        self._keep = []

    def __del__(self):
        b = getattr(self, "_batch", None)
        if b:
            try:
                lib().nrocm_batch_destroy(b)
            except Exception:
                pass

    def set_graph(self, enable: bool) -> None:
        check("nrocm_batch_set_graph", lib().nrocm_batch_set_graph(self._batch, int(enable)))

    def autotune(self, mode: int = EXACT) -> dict:
        """Measure and keep the fastest lanes-per-matrix schedule of every size
        bin (nrocm_batch_autotune).  Returns {padded size: lanes per matrix}."""
        lanes = (C.c_int * 8)()
        h = handle()
        check("nrocm_batch_autotune", lib().nrocm_batch_autotune(h, self._batch, mode, lanes), h)
        return {8 * (c + 1): lanes[c] for c in range(8) if lanes[c]}

    def set_schedule(self, lanes: dict, mode: int = EXACT) -> None:
        """lanes: {padded size: lanes per matrix} (nrocm_batch_set_schedule)."""
        arr = (C.c_int * 8)()
        for nb, g in lanes.items():
            arr[nb // 8 - 1] = int(g)
        check("nrocm_batch_set_schedule", lib().nrocm_batch_set_schedule(self._batch, mode, arr))

    def schedule(self, mode: int = EXACT) -> dict:
        lanes = (C.c_int * 8)()
        check("nrocm_batch_schedule", lib().nrocm_batch_schedule(self._batch, mode, lanes))
        return {8 * (c + 1): lanes[c] for c in range(8) if lanes[c]}

    # -------------------------------------------------------------- packed storage
    def alloc(self):
        A = torch.empty(self.sizeA, dtype=torch.float64, device=self.device)
        B = torch.empty(self.sizeB, dtype=torch.float64, device=self.device)
        return A, B

    def pointers(self, A: torch.Tensor, B: torch.Tensor):
        return _pointer_array(A, self.offA), _pointer_array(B, self.offB)

    def solve(self, A: torch.Tensor, B: torch.Tensor, mode: int = EXACT, ptrs=None) -> torch.Tensor:
        """Factor every A_j in place and overwrite B_j with A_j^{-1} B_j.
        Returns the per-member info (device int32)."""
        if ptrs is None:
            ptrs = self.pointers(A, B)
        pa, pb = ptrs
        info = torch.empty(self.count, dtype=torch.int32, device=self.device)
        self.solve_ptrs(pa, pb, info, mode)
        return info

    def solve_ptrs(self, pa: torch.Tensor, pb: torch.Tensor, info: torch.Tensor, mode: int = EXACT):
        h = handle()
        st = lib().nrocm_dpotrf_potrs_batched(h, self._batch, _ptr(pa), _ptr(pb), _ptr(info), mode)
        check("nrocm_dpotrf_potrs_batched", st, h)

    def backward(self, L: torch.Tensor, X: torch.Tensor, dX: torch.Tensor):
        """Adjoint for packed storage: returns (dA packed n_j x n_j, dB packed)."""
        dA = torch.empty(self.sizeA, dtype=torch.float64, device=self.device)
        dB = torch.empty(self.sizeB, dtype=torch.float64, device=self.device)
        ptrs = [_pointer_array(L, self.offA), _pointer_array(X, self.offB),
                _pointer_array(dX, self.offB), _pointer_array(dA, self.offA),
                _pointer_array(dB, self.offB)]
        self.backward_raw(*ptrs)
        return dA, dB

    def backward_ptrs(self, Ls, Xs, dXs, dAs, dBs):
        self.backward_raw(*[_pointer_list(v, self.device) for v in (Ls, Xs, dXs, dAs, dBs)])

    def backward_raw(self, pL, pX, pdX, pdA, pdB):
        h = handle()
        st = lib().nrocm_dpotrs_backward_batched(h, self._batch, _ptr(pL), _ptr(pX), _ptr(pdX),
                                                 _ptr(pdA), _ptr(pdB))
        check("nrocm_dpotrs_backward_batched", st, h)


class LUBatch:
    """count general systems A_j (n_j x n_j, n_j <= 64) with B_j (n_j x nrhs):
    LU::new + LU::solve_mut per member (nrocm_dgetrf_getrs_batched); pivots
    packed back to back (member j's n_j pivots at offP[j])."""

    def __init__(self, ns: Sequence[int], nrhs: int, device="cuda"):
        self.ns = np.ascontiguousarray(ns, dtype=np.int32)
        self.count = len(self.ns)
        self.nrhs = int(nrhs)
        self.device = torch.device(device)
        n64 = self.ns.astype(np.int64)
        self.offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
        self.offB = np.concatenate([[0], np.cumsum(n64 * self.nrhs)[:-1]]).astype(np.int64)
        self.offP = np.concatenate([[0], np.cumsum(n64)[:-1]]).astype(np.int64)
        self.sizeA = int((n64 * n64).sum())
        self.sizeB = int((n64 * self.nrhs).sum())
        self.sizeP = int(n64.sum())
        self._h = handle()
        bp = C.c_void_p()
        check("nrocm_batch_create",
              lib().nrocm_batch_create(self._h, self.count,
                                       self.ns.ctypes.data_as(C.POINTER(C.c_int)), self.nrhs,
                                       C.byref(bp)), self._h)
        self._batch = bp.value

    def __del__(self):
        b = getattr(self, "_batch", None)
        if b:
            try:
                lib().nrocm_batch_destroy(b)
            except Exception:
                pass

    def pointers(self, A: torch.Tensor, B: torch.Tensor, P: torch.Tensor):
        return _pointer_array(A, self.offA), _pointer_array(B, self.offB), _pointer_array(P, self.offP)

    def solve(self, A: torch.Tensor, B: torch.Tensor, mode: int = EXACT):
        """Factor every A_j in place (packed L \\ U) and overwrite B_j with
        A_j^{-1} B_j.  Returns (pivots packed int32, info int32)."""
        P = torch.empty(max(self.sizeP, 1), dtype=torch.int32, device=self.device)
        info = torch.empty(self.count, dtype=torch.int32, device=self.device)
        pa, pb, pp = self.pointers(A, B, P)
        self.solve_ptrs(pa, pp, pb, info, mode)
        return P[:self.sizeP], info

    def backward(self, LU: torch.Tensor, P: torch.Tensor, X: torch.Tensor, dX: torch.Tensor):
        """Adjoint of X = A^{-1} B from the packed factors and pivots that solve()
        left: returns (dA packed n_j x n_j, dB packed) for the upstream dX."""
        dA = torch.empty(max(self.sizeA, 1), dtype=torch.float64, device=self.device)
        dB = torch.empty(max(self.sizeB, 1), dtype=torch.float64, device=self.device)
        ptrs = [_pointer_array(LU, self.offA), _pointer_array(P, self.offP), _pointer_array(X, self.offB),
                _pointer_array(dX, self.offB), _pointer_array(dA, self.offA), _pointer_array(dB, self.offB)]
        h = handle()
        st = lib().nrocm_dgetrs_backward_batched(h, self._batch, *[_ptr(t) for t in ptrs])
        check("nrocm_dgetrs_backward_batched", st, h)
        return dA[:self.sizeA], dB[:self.sizeB]

    def solve_ptrs(self, pa, pp, pb, info, mode: int = EXACT):
        h = handle()
        st = lib().nrocm_dgetrf_getrs_batched(h, self._batch, _ptr(pa), _ptr(pp) if pp is not None else None,
                                              _ptr(pb), _ptr(info), mode)
        check("nrocm_dgetrf_getrs_batched", st, h)


class LeastSquaresBatch:
    """count tall f32 systems A_j (m_j x n) with b_j (m_j), solved by blocked
    Householder QR (FP32 MFMA trailing updates) plus FP64 refinement."""

    def __init__(self, ms: Sequence[int], n: int, device="cuda"):
        self.ms = np.ascontiguousarray(ms, dtype=np.int32)
        self.count = len(self.ms)
        self.n = int(n)
        self.device = torch.device(device)
        m64 = self.ms.astype(np.int64)
        self.offA = np.concatenate([[0], np.cumsum(m64 * self.n)[:-1]]).astype(np.int64)
        self.offb = np.concatenate([[0], np.cumsum(m64)[:-1]]).astype(np.int64)
        self.sizeA = int((m64 * self.n).sum())
        self.sizeb = int(m64.sum())
        self._h = handle()
        bp = C.c_void_p()
        check("nrocm_batch_create",
              lib().nrocm_batch_create(self._h, self.count,
                                       self.ms.ctypes.data_as(C.POINTER(C.c_int)), 1, C.byref(bp)),
              self._h)
        self._batch = bp.value

    def __del__(self):
        b = getattr(self, "_batch", None)
        if b:
            try:
                lib().nrocm_batch_destroy(b)
            except Exception:
                pass

    def pointers(self, A: torch.Tensor, b: torch.Tensor):
        return _pointer_array(A, self.offA), _pointer_array(b, self.offb)

    def solve(self, A: torch.Tensor, b: torch.Tensor, refine: int = 1, ptrs=None):
        """Returns (x: count x n float64, info int32); A and b are not modified."""
        if ptrs is None:
            ptrs = self.pointers(A, b)
        x = torch.empty(self.count * self.n, dtype=torch.float64, device=self.device)
        info = torch.empty(self.count, dtype=torch.int32, device=self.device)
        self.solve_ptrs(ptrs[0], ptrs[1], x, info, refine)
        return x.view(self.count, self.n), info

    def solve_ptrs(self, pa, pb, x, info, refine: int = 1):
        h = handle()
        st = lib().nrocm_sgels_batched(h, self._batch, self.n, _ptr(pa), _ptr(pb), _ptr(x), refine,
                                       _ptr(info))
        check("nrocm_sgels_batched", st, h)
