"""Hierarchical matrix on the GPU -- mirror of src/solve/h_matrix.rs.

HMatrix.new_from(predata, side, leaf_cardinality=32, tolerance=1e-4, eta=4.0)
                                          h_matrix.rs:127-148 (solve.rs:86-92,114-118)
HMatrix.new_from_matrix(a, coords, leaf_cardinality, tolerance, eta=4.0)
                                          h_matrix.rs:127-148 with the closure of an
                                          explicit matrix (h_matrix.rs:282-298)
HMatrix.gemv(alpha, x, beta, b)           h_matrix.rs:218-234
HMatrix.to_full()                         h_matrix.rs:236-247
HMatrix.get_norm(), get_num_eqn()         h_matrix.rs:149-150
HMatrix.stats()                           h_matrix.rs:197-216 (print_stats)

The blocks are built and stored on the device by libccab (batched ACA with
SVD recompression for the admissible blocks, dense near-field blocks).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import capi
from .preprocess import PreData


def _ptr(t) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr()) if t is not None else None


class HMatrix:
    def __init__(self, handle: ct.c_void_p, n: int, device: torch.device, ctx: capi.Context,
                 ncol: int | None = None):
        self._h = handle
        self.num_eqn = n
        self.num_col = n if ncol is None else ncol
        self.device = device
        self.ctx = ctx

    @classmethod
    def new_from_matrix(cls, a: torch.Tensor, coords, leaf_cardinality: int = 32,
                        tolerance: float = 1e-4, eta: float = 4.0, ctx=None,
                        storage: int = capi.C128) -> "HMatrix":
        """H matrix of the explicit n x n complex128 device matrix a (any
        strides; copied to column-major), clustered by the points coords
        (n x 3, host): HMatrix::new_from with the get_row_or_column closure of
        a matrix, as the reference's h_matrix.rs tests build it."""
        n = a.shape[0]
        if a.dim() != 2 or a.shape[1] != n:
            raise ValueError("the operator must be square")
        xyz = np.ascontiguousarray(np.asarray(coords, dtype=np.float64).reshape(n, -1)[:, :3])
        ctx = ctx or capi.default_context(a.device.index or 0)
        acm = a.to(torch.complex128).t().contiguous()  # column-major storage
        h = ct.c_void_p()
        ctx.call("ccab_hmatrix_build_matrix", n, _ptr(acm), n,
                 xyz.ctypes.data_as(ct.c_void_p), int(leaf_cardinality), float(eta),
                 float(tolerance), int(storage), ct.byref(h))
        ctx.synchronize()
        return cls(h, n, a.device, ctx)

    @classmethod
    def new_from(cls, predata: PreData, side: int = capi.LHS, leaf_cardinality: int = 32,
                 tolerance: float = 1e-4, eta: float = 4.0, ctx=None,
                 storage: int = capi.C128) -> "HMatrix":
        """H matrix of the LHS (a H + b G) or RHS side matrix at the current frequency.
        storage C64 keeps the blocks (built in FP64) in complex FP32."""
        ctx = ctx or capi.default_context(predata.device.index or 0)
        h = ct.c_void_p()
        ctx.call("ccab_hmatrix_build_ex", ct.byref(predata.physics()), _ptr(predata.d_cpts),
                 predata.get_num_eqn(), int(side), int(leaf_cardinality), float(eta),
                 float(tolerance), int(storage), ct.byref(h))
        return cls(h, predata.get_num_eqn(), predata.device, ctx)

    def get_num_eqn(self) -> int:
        return self.num_eqn

    def get_norm(self) -> float:
        return self.stats()["norm"]

    def stats(self) -> dict:
        st = (ct.c_int64 * 12)()
        nrm = ct.c_double()
        rc = capi.lib().ccab_hmatrix_stats_ex(self._h, st, 12, ct.byref(nrm))
        if rc != capi.OK:
            raise capi.CcabError(rc, "ccab_hmatrix_stats_ex failed")
        n2 = float(self.num_eqn) * float(self.num_col)
        return {"admissible": st[0], "inadmissible": st[1], "storage": st[2], "rank_sum": st[3],
                "dense_fallback": st[4], "norm": nrm.value,
                "admissible_ratio": 100.0 * st[0] / max(st[0] + st[1], 1),
                "compression_ratio": 100.0 * (1.0 - st[2] / n2),
                # admissible blocks by ACA rank before recompression
                "aca_rank_buckets": {"<=8": st[5], "9-16": st[6], "17-32": st[7], "33-64": st[8]},
                "max_rank": st[9]}

    def gemv(self, alpha, x: torch.Tensor, beta, b: torch.Tensor) -> torch.Tensor:
        """b = alpha * H x + beta * b (in place on b, complex128 device vectors)."""
        if x.dtype != torch.complex128 or b.dtype != torch.complex128:
            raise TypeError("gemv works on complex128 vectors")
        if x.numel() != self.num_col or b.numel() != self.num_eqn:
            raise ValueError("dimension mismatch in H matrix gemv")
        a = (ct.c_double * 2)(complex(alpha).real, complex(alpha).imag)
        bb = (ct.c_double * 2)(complex(beta).real, complex(beta).imag)
        self.ctx.call("ccab_hmatrix_gemv", self._h, a, _ptr(x.contiguous()), bb, _ptr(b))
        self.ctx.synchronize()
        return b

    def to_full(self) -> torch.Tensor:
        """The n x n matrix (column-major view) the H matrix stands for."""
        m, n = self.num_eqn, self.num_col
        a = torch.empty((n, m), dtype=torch.complex128, device=self.device)
        self.ctx.call("ccab_hmatrix_to_full", self._h, _ptr(a), m)
        self.ctx.synchronize()
        return a.t()  # storage is column-major

    def close(self):
        if self._h:
            capi.lib().ccab_hmatrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
