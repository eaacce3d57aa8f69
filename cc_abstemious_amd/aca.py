"""Adaptive cross approximation on the GPU -- mirror of src/solve/aca.rs.

ACA(tol, a)              ACA::new(tol, m, n, get_row, get_column)   aca.rs:22-33
                         with the closures of an explicit matrix a (aca.rs:303-304)
.gemv(alpha, x, beta, b) ACA::gemv                                  aca.rs:181-195
.get_norm()              ACA::get_norm (Frobenius norm estimate)    aca.rs:212-241
.get_num_uv()            ACA::get_num_uv (rank after recompression) aca.rs:50
.to_full()               the m x n matrix U V^T

libccab runs the decomposition as one admissible block of its batched ACA
kernel (partial pivoting, the reference's seeded reference rows, SVD
recompression); the handle is an m x n ccab_hmatrix.
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import capi
from .h_matrix import HMatrix, _ptr


class ACA(HMatrix):
    def __init__(self, tol: float, a: torch.Tensor, ctx=None):
        if a.dim() != 2:
            raise ValueError("ACA takes a matrix")
        m, n = a.shape
        ctx = ctx or capi.default_context(a.device.index or 0)
        acm = a.to(torch.complex128).t().contiguous()  # column-major storage
        h = ct.c_void_p()
        ctx.call("ccab_aca_build", int(m), int(n), _ptr(acm), int(m), float(tol), ct.byref(h))
        ctx.synchronize()
        super().__init__(h, m, a.device, ctx, ncol=n)

    def get_num_uv(self) -> int:
        st = self.stats()
        if st["dense_fallback"]:
            raise RuntimeError("the ACA did not converge within 64 terms (stored dense)")
        return st["rank_sum"]
