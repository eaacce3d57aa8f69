"""Incident wave on the surface -- mirror of get_incident_surface
(src/incident_wave.rs:7-63), evaluated on the GPU."""
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
