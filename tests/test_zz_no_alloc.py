"""ccab_field_noalloc (scratch in the library's own device array, no
allocation) against ccab_field (context buffers) through the C ABI: two calls
in a row, the second on a smaller input, must give the same field values,
bit for bit (same kernels, same chunking at these sizes); inputs beyond the
scratch are rejected by the host before any launch."""
import ctypes as ct
import math

import numpy as np
import pytest
import torch

from cc_abstemious_amd import capi, mesh
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ct.c_void_p(t.data_ptr()) if t is not None else None


def _case(level, nfreq, nfp, seed, kind):
    rng = np.random.default_rng(seed)
    m = mesh.icosphere(level)
    fs = list(np.linspace(100.0, 400.0, nfreq))
    w = IncidentWave(kind, tuple(rng.standard_normal(3)), (1.0, 0.3))
    pd = PreData(m, fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY,
                 (0.0, 0.0), w, Solver())
    n = pd.get_num_eqn()
    fp = rng.standard_normal((nfp, 3))
    fp = 3.0 * fp / np.linalg.norm(fp, axis=1)[:, None]
    d_fps = torch.from_numpy(np.ascontiguousarray(fp)).cuda()
    phi = torch.from_numpy(rng.standard_normal((nfreq, n)) + 1j * rng.standard_normal((nfreq, n))).cuda()
    vn = torch.from_numpy(rng.standard_normal((nfreq, n)) + 1j * rng.standard_normal((nfreq, n))).cuda()
    return pd, fs, w.to_c(), d_fps, phi, vn


def _call(ctx, name, pd, fs, w, d_fps, phi, vn, total, pressure):
    nfreq, n = phi.shape
    nfp = d_fps.shape[0]
    out = torch.full((nfreq, nfp), complex("nan"), dtype=torch.complex128, device="cuda")
    inc = torch.full((nfreq, nfp), complex("nan"), dtype=torch.complex128, device="cuda")
    ctx.call(name, pd.physics_array(fs), nfreq, ct.byref(w), _ptr(pd.d_cpts), n, _ptr(d_fps), nfp,
             _ptr(phi), _ptr(vn), total, pressure, _ptr(out), _ptr(inc))
    ctx.synchronize()
    return out.cpu().numpy(), inc.cpu().numpy()


def test_field_noalloc_matches_field_twice():
    ctx = capi.default_context(0)
    for label, args, total, pressure in (
            ("level 4, 8 frequencies, 300 field points, plane", (4, 8, 300, 1, capi.WAVE_PLANE), 1, 1),
            ("level 2, 3 frequencies, 70 field points, spherical", (2, 3, 70, 2, capi.WAVE_SPHERICAL), 0, 0)):
        print(f"[case] ccab_field_noalloc vs ccab_field: {label}", flush=True)
        case = _case(*args)
        a_out, a_inc = _call(ctx, "ccab_field", *case, total, pressure)
        b_out, b_inc = _call(ctx, "ccab_field_noalloc", *case, total, pressure)
        assert np.isfinite(b_out).all() and np.isfinite(b_inc).all()
        np.testing.assert_array_equal(b_out, a_out)
        np.testing.assert_array_equal(b_inc, a_inc)


def test_field_noalloc_rejects_oversized_inputs():
    ctx = capi.default_context(0)
    pd, fs, w, d_fps, phi, vn = _case(1, 1, 4, 3, capi.WAVE_PLANE)
    nfreq, n = phi.shape
    out = torch.empty((1, 4), dtype=torch.complex128, device="cuda")
    phys = pd.physics_array(fs)
    # more frequencies than the limit: rejected on the host (nothing launched,
    # no buffer read)
    big = (capi.Physics * 300)(*([phys[0]] * 300))
    with pytest.raises(capi.CcabError, match="at most 256 frequencies"):
        ctx.call("ccab_field_noalloc", big, 300, ct.byref(w), _ptr(pd.d_cpts), n, _ptr(d_fps), 4,
                 _ptr(phi), _ptr(vn), 0, 0, _ptr(out), None)
    # a surface whose point copy alone exceeds the 32 MiB scratch: rejected on
    # the host before any launch (n is only compared, nothing is read)
    with pytest.raises(capi.CcabError, match="scratch"):
        ctx.call("ccab_field_noalloc", phys, 1, ct.byref(w), _ptr(pd.d_cpts), 1 << 20,
                 _ptr(d_fps), 4, _ptr(phi), _ptr(vn), 0, 0, _ptr(out), None)
