"""Incident wave on the GPU -- mirror of src/incident_wave.rs:
get_incident_surface (7-63) and get_incident_field (66-100)."""
from __future__ import annotations

import ctypes as ct

import torch

from . import capi
from .preprocess import PreData


def get_incident_surface(predata: PreData, ctx=None) -> torch.Tensor:
    ctx = ctx or capi.default_context(predata.device.index or 0)
    n = predata.get_num_eqn()
    out = torch.empty(n, dtype=torch.complex128, device=predata.device)
    w = predata.incident_wave.to_c()
    ctx.call("ccab_incident_surface", ct.byref(predata.physics()), ct.byref(w),
             ct.c_void_p(predata.d_cpts.data_ptr()), n, ct.c_void_p(out.data_ptr()))
    ctx.synchronize()
    return out


def get_incident_field(predata: PreData, ctx=None) -> torch.Tensor:
    """Incident potential at predata's field points (nfp,)."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    if predata.d_fps is None or predata.d_fps.shape[0] == 0:
        raise ValueError("no field points: call predata.set_field_points")
    nfp = predata.d_fps.shape[0]
    out = torch.empty(nfp, dtype=torch.complex128, device=predata.device)
    w = predata.incident_wave.to_c()
    ctx.call("ccab_incident_field", ct.byref(predata.physics()), ct.byref(w),
             ct.c_void_p(predata.d_fps.data_ptr()), nfp, ct.c_void_p(out.data_ptr()))
    ctx.synchronize()
    return out
