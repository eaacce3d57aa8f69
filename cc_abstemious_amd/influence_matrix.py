"""Influence matrices on the GPU -- mirror of src/influence_matrix.rs.

get_dense_surface_matrices   influence_matrix.rs:64-98
get_dense_field_matrices     influence_matrix.rs:101-127
get_surface_row_or_column    influence_matrix.rs:170-231
assemble_system              fused assembly + BC algebra of solve.rs:24-71
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import capi
from .preprocess import PreData, colmajor

LHS, RHS = capi.LHS, capi.RHS


def _ptr(t: torch.Tensor) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr())


def get_dense_surface_matrices(predata: PreData, prec: int = capi.C128, ctx=None):
    """Returns (H, G) as n x n column-major cuda tensors (complex128 or complex64)."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    n = predata.get_num_eqn()
    dt = torch.complex128 if prec == capi.C128 else torch.complex64
    H = colmajor(n, n, dt, predata.device)
    G = colmajor(n, n, dt, predata.device)
    ph = predata.physics()
    ctx.call("ccab_dense_surface_matrices", ct.byref(ph), _ptr(predata.d_cpts), n, prec, _ptr(H),
             _ptr(G), n)
    ctx.synchronize()
    return H, G


def get_dense_field_matrices(predata: PreData, ctx=None):
    """Returns (M, L) as nfp x n column-major complex128 cuda tensors."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    if predata.d_fps is None or predata.d_fps.shape[0] == 0:
        raise ValueError("no field points: call predata.set_field_points")
    n, nfp = predata.get_num_eqn(), predata.d_fps.shape[0]
    M = colmajor(nfp, n, torch.complex128, predata.device)
    L = colmajor(nfp, n, torch.complex128, predata.device)
    ctx.call("ccab_dense_field_matrices", ct.byref(predata.physics()), _ptr(predata.d_cpts), n,
             _ptr(predata.d_fps), nfp, _ptr(M), _ptr(L), nfp)
    ctx.synchronize()
    return M, L


def get_surface_row_or_column(predata: PreData, i, j, side: int = LHS, ctx=None) -> torch.Tensor:
    """One row (len(i) == 1) or column (len(j) == 1) of alpha*H + beta*G."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    ii = torch.as_tensor(list(i) if not isinstance(i, torch.Tensor) else i, dtype=torch.int64,
                         device=predata.device).reshape(-1)
    jj = torch.as_tensor(list(j) if not isinstance(j, torch.Tensor) else j, dtype=torch.int64,
                         device=predata.device).reshape(-1)
    out = torch.empty(max(ii.numel(), jj.numel()), dtype=torch.complex128, device=predata.device)
    ctx.call("ccab_surface_row_or_column", ct.byref(predata.physics()), _ptr(predata.d_cpts),
             predata.get_num_eqn(), side, _ptr(ii), ii.numel(), _ptr(jj), jj.numel(), _ptr(out))
    ctx.synchronize()
    return out


def assemble_system(predata: PreData, rhs_inc: torch.Tensor, frequencies=None, want_c128=True,
                    want_c64=False, ctx=None):
    """Fused LHS (alpha H + beta G) + right-hand side for one or many frequencies.
    Returns (lhs128 or None, lhs64 or None, rhs) with lhs of shape (batch, n, n)
    column-major per system and rhs (batch, n)."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    fs = [predata.get_frequency()] if frequencies is None else list(frequencies)
    nb = len(fs)
    n = predata.get_num_eqn()
    dev = predata.device
    l128 = torch.empty((nb, n, n), dtype=torch.complex128, device=dev) if want_c128 else None
    l64 = torch.empty((nb, n, n), dtype=torch.complex64, device=dev) if want_c64 else None
    rhs = torch.empty((nb, n), dtype=torch.complex128, device=dev)
    ri = rhs_inc.reshape(nb, n).contiguous()
    ph = predata.physics_array(fs)
    ctx.call("ccab_assemble_system", ph, nb, _ptr(predata.d_cpts), n,
             _ptr(l128) if l128 is not None else None, _ptr(l64) if l64 is not None else None, n,
             n * n, _ptr(ri), _ptr(rhs))
    ctx.synchronize()
    # storage of system b is column-major: element (i, j) at i + j n -> view as [b, j, i]
    t = lambda x: None if x is None else x.transpose(1, 2)
    return t(l128), t(l64), rhs
