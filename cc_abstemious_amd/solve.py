"""Surface solve on the GPU -- mirror of src/solve.rs and src/solve/gmres.rs.

solve_for_surface(predata, rhs_inc) -> (phi, vn)      solve.rs:11-71
solve_lu(a, x)                                        solve.rs:166-171
GMRES(max_it, thresh); .a = matrix or .hmatrix = HMatrix; .solve(x)
                                                      gmres.rs:22-59
sweep(predata) -> (phi, vn, stats)                    analysis.rs:86-111 (surface part)
solve_bodies([predata, ...]) -> (phis, vns, stats)    solve.rs:11-71 over uneven bodies
get_field(predata, m, l, phi, vn, phi_inc_fp)         solve.rs:173-185
field(predata, phi, vn) -> (scattered, incident)      analysis.rs:100-104 + postprocess.rs:70-88,
                                                      fused, batched over frequencies
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import capi
from .preprocess import PreData


def _ptr(t) -> ct.c_void_p:
    return ct.c_void_p(t.data_ptr()) if t is not None else None


def _colmajor_ok(a: torch.Tensor) -> int:
    """Leading dimension of a column-major 2-D tensor view, or raise."""
    if a.dim() != 2 or a.stride(0) != 1 or a.stride(1) < a.shape[0]:
        raise ValueError("matrix must be a column-major view (stride (1, ld))")
    return a.stride(1)


def solve_for_surface(predata: PreData, rhs_inc: torch.Tensor, ctx=None):
    """Assemble and solve the surface system at the current frequency."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    n = predata.get_num_eqn()
    phi = torch.empty(n, dtype=torch.complex128, device=predata.device)
    vn = torch.empty(n, dtype=torch.complex128, device=predata.device)
    st = capi.SolveStats()
    cfg = predata.solver.to_c()
    ri = rhs_inc.to(torch.complex128).contiguous()
    ctx.call("ccab_solve_surface", ct.byref(predata.physics()), _ptr(predata.d_cpts), n, _ptr(ri),
             ct.byref(cfg), _ptr(phi), _ptr(vn), ct.byref(st))
    ctx.synchronize()
    predata.last_stats = st.as_dict()
    return phi, vn


def sweep(predata: PreData, frequencies=None, x0: torch.Tensor | None = None, ctx=None):
    """Surface solution at every frequency: phi, vn of shape (nfreq, n).
    x0 (nfreq, n): initial guess of the solved unknown (ITERATIVE_MIXED only)."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    fs = predata.get_frequencies() if frequencies is None else list(frequencies)
    n = predata.get_num_eqn()
    phi = torch.empty((len(fs), n), dtype=torch.complex128, device=predata.device)
    vn = torch.empty((len(fs), n), dtype=torch.complex128, device=predata.device)
    stats = (capi.SolveStats * len(fs))()
    cfg = predata.solver.to_c()
    w = predata.incident_wave.to_c()
    if x0 is None:
        ctx.call("ccab_sweep", predata.physics_array(fs), len(fs), ct.byref(w),
                 _ptr(predata.d_cpts), n, ct.byref(cfg), _ptr(phi), _ptr(vn), stats)
    else:
        x0 = x0.to(torch.complex128).contiguous()
        if x0.shape != (len(fs), n):
            raise ValueError("x0 must have shape (nfreq, n)")
        ctx.call("ccab_sweep_warm", predata.physics_array(fs), len(fs), ct.byref(w),
                 _ptr(predata.d_cpts), n, ct.byref(cfg), _ptr(x0), _ptr(phi), _ptr(vn), stats)
    ctx.synchronize()
    return phi, vn, [s.as_dict() for s in stats]


def solve_bodies(bodies, x0=None, solver=None, ctx=None):
    """solve_for_surface (solve.rs:11-71) for several bodies of uneven size in
    one batched call: bodies is a list of PreData, each solved at its current
    frequency with its own incident wave (one BC type for all); solver
    defaults to the first body's.  x0: optional list of initial guesses (one
    (n_b,) tensor per body; ITERATIVE_MIXED or HIERARCHICAL).  Returns the
    lists phi, vn (one (n_b,) tensor per body) and the per-body stats."""
    if not bodies:
        raise ValueError("no bodies")
    ctx = ctx or capi.default_context(bodies[0].device.index or 0)
    nb = len(bodies)
    ns = [pd.get_num_eqn() for pd in bodies]
    n_arr = (ct.c_int64 * nb)(*ns)
    cpts = torch.cat([pd.d_cpts for pd in bodies]).contiguous()
    ph = (capi.Physics * nb)()
    w = (capi.Wave * nb)()
    for i, pd in enumerate(bodies):
        ph[i] = pd.physics()
        w[i] = pd.incident_wave.to_c()
    cfg = (solver or bodies[0].solver).to_c()
    total = sum(ns)
    dev = bodies[0].device
    phi = torch.empty(total, dtype=torch.complex128, device=dev)
    vn = torch.empty(total, dtype=torch.complex128, device=dev)
    x0p = None
    if x0 is not None:
        if len(x0) != nb or any(x.numel() != n for x, n in zip(x0, ns)):
            raise ValueError("x0 must hold one (n_b,) guess per body")
        x0p = torch.cat([x.reshape(-1).to(dev) for x in x0]).to(torch.complex128).contiguous()
    stats = (capi.SolveStats * nb)()
    ctx.call("ccab_solve_bodies", nb, n_arr, _ptr(cpts), ph, w, ct.byref(cfg), _ptr(x0p),
             _ptr(phi), _ptr(vn), stats)
    ctx.synchronize()
    return list(phi.split(ns)), list(vn.split(ns)), [s.as_dict() for s in stats]


def gmres_ir_batched(A128: torch.Tensor | None, A64: torch.Tensor, x: torch.Tensor,
                     x0: torch.Tensor | None = None, tol=1e-10, inner_tol=1e-4, restart=50,
                     max_cycles=20, ref_mgs=False, ctx=None):
    """Batched GMRES on the c64 operator with FP64 refinement against A128
    (column-major systems, see _check_batch); x (batch, n): b -> x."""
    ctx = ctx or capi.default_context(A64.device.index or 0)
    nb, n = _check_batch(A64)
    if A128 is not None:
        _check_batch(A128)
    cfg = capi.solver_cfg(capi.SOLVER_ITERATIVE_MIXED,
                          capi.gmres_params(n, tol, restart, ref_mgs, capi.STOP_RELATIVE_RESIDUAL,
                                            capi.C64), inner_tol, max_cycles)
    stats = (capi.SolveStats * nb)()
    ctx.call("ccab_gmres_ir_batched", n, nb, _ptr(A128), _ptr(A64), n, n * n, _ptr(x), _ptr(x0),
             ct.byref(cfg), stats)
    ctx.synchronize()
    return x, [s.as_dict() for s in stats]


def solve_lu(a: torch.Tensor, x: torch.Tensor, ctx=None):
    """x <- a^{-1} x (a is not modified: the library factors a copy)."""
    ctx = ctx or capi.default_context(a.device.index or 0)
    lda = _colmajor_ok(a)
    if a.dtype != torch.complex128 or x.dtype != torch.complex128 or not x.is_contiguous():
        raise ValueError("solve_lu takes complex128 tensors and a contiguous x")
    ctx.call("ccab_solve_lu", a.shape[0], _ptr(a), lda, _ptr(x))
    ctx.synchronize()
    return x


def _check_batch(A: torch.Tensor):
    """A (batch, n, n) holds column-major systems back to back: A[b, i, j]
    at offset b n^2 + i + j n, i.e. A = S.transpose(1, 2) for contiguous S."""
    nb, n, m = A.shape
    if n != m or A.stride(1) != 1 or A.stride(2) != n or (nb > 1 and A.stride(0) != n * n):
        raise ValueError("expected column-major systems packed back to back "
                         "(use S.transpose(1, 2) of a contiguous (batch, n, n) S)")
    return nb, n


def lu_factor_batched(A: torch.Tensor, ctx=None):
    """In-place batched LU of A (batch, n, n), column-major systems;
    returns ipiv (batch, n) int32 (0-based, LAPACK order)."""
    ctx = ctx or capi.default_context(A.device.index or 0)
    nb, n = _check_batch(A)
    prec = capi.C128 if A.dtype == torch.complex128 else capi.C64
    ipiv = torch.empty((nb, n), dtype=torch.int32, device=A.device)
    ctx.call("ccab_lu_factor_batched", prec, n, nb, _ptr(A), n, n * n, _ptr(ipiv))
    ctx.synchronize()
    return ipiv


def lu_solve_batched(LU: torch.Tensor, ipiv: torch.Tensor, B: torch.Tensor, ctx=None):
    """B (batch, n) or (batch, nrhs, n) overwritten with the solutions."""
    ctx = ctx or capi.default_context(LU.device.index or 0)
    nb, n = _check_batch(LU)
    prec = capi.C128 if LU.dtype == torch.complex128 else capi.C64
    nrhs = 1 if B.dim() == 2 else B.shape[1]
    if not B.is_contiguous():
        raise ValueError("B must be contiguous")
    ctx.call("ccab_lu_solve_batched", prec, n, nb, _ptr(LU), n, n * n, _ptr(ipiv), _ptr(B), n,
             n * nrhs, nrhs)
    ctx.synchronize()
    return B


def solve_mixed_batched(A128: torch.Tensor, A64: torch.Tensor, x: torch.Tensor, ir_tol=1e-12,
                        ir_max_steps=10, ctx=None):
    """c64 LU (destroys A64) + FP64 refinement against A128; x (batch, n) b -> x."""
    ctx = ctx or capi.default_context(A128.device.index or 0)
    nb, n = _check_batch(A128)
    _check_batch(A64)
    if A128.dtype != torch.complex128 or A64.dtype != torch.complex64 or not x.is_contiguous():
        raise ValueError("A128 complex128, A64 complex64, x contiguous complex128")
    stats = (capi.SolveStats * nb)()
    ctx.call("ccab_solve_mixed_batched", n, nb, _ptr(A128), _ptr(A64), n, n * n, _ptr(x),
             float(ir_tol), int(ir_max_steps), stats)
    ctx.synchronize()
    return x, [s.as_dict() for s in stats]


class GMRES:
    """Restarted GMRES with the reference's parameters (gmres.rs:12-59)."""

    def __init__(self, max_it: int, thresh: float, restart: int = 0, ref_mgs: bool = True,
                 stop_rule: int = capi.STOP_BACKWARD_ERROR):
        self.max_it = max_it
        self.thresh = thresh
        self.restart = restart
        self.ref_mgs = ref_mgs
        self.stop_rule = stop_rule
        self.a = None
        self.hmatrix = None          # gmres.rs:18 (an HMatrix operator instead of .a)
        self.num_mv = 0
        self.stats = None

    def solve(self, x: torch.Tensor, x0: torch.Tensor | None = None, ctx=None):
        """x: b on entry, solution on exit (complex128, contiguous)."""
        if self.hmatrix is not None:
            h = self.hmatrix
            ctx = ctx or h.ctx
            p = capi.gmres_params(self.max_it, self.thresh, self.restart, self.ref_mgs,
                                  self.stop_rule, capi.C128)
            st = capi.SolveStats()
            ctx.call("ccab_gmres_hmatrix", h._h, _ptr(x), _ptr(x0) if x0 is not None else None,
                     ct.byref(p), ct.byref(st))
            ctx.synchronize()
            self.stats = st.as_dict()
            self.num_mv = st.num_mv
            return x
        if self.a is None:
            raise ValueError("GMRES needs an operator: set .a or .hmatrix")
        a = self.a
        ctx = ctx or capi.default_context(a.device.index or 0)
        lda = _colmajor_ok(a)
        prec = capi.C128 if a.dtype == torch.complex128 else capi.C64
        p = capi.gmres_params(self.max_it, self.thresh, self.restart, self.ref_mgs,
                              self.stop_rule, prec)
        st = capi.SolveStats()
        ctx.call("ccab_gmres_dense", a.shape[0], _ptr(a), lda, _ptr(x),
                 _ptr(x0) if x0 is not None else None, ct.byref(p), ct.byref(st))
        ctx.synchronize()
        self.stats = st.as_dict()
        self.num_mv = st.num_mv
        return x


def get_field(predata: PreData, m: torch.Tensor, l: torch.Tensor, phi: torch.Tensor,
              vn: torch.Tensor, phi_inc_fp: torch.Tensor) -> torch.Tensor:
    """phi_fp = (total ? phi_inc_fp : 0) - M phi + L vn (solve.rs:173-185), from
    explicit M, L (see influence_matrix.get_dense_field_matrices); the
    scattered part changes sign with capi.FIX_FIELD_SIGN."""
    out = phi_inc_fp.clone() if predata.output_total else torch.zeros_like(phi_inc_fp)
    if getattr(predata, "corrections", 0) & capi.FIX_FIELD_SIGN:
        return out + m @ phi - l @ vn
    return out - m @ phi + l @ vn


def field(predata: PreData, phi: torch.Tensor, vn: torch.Tensor, frequencies=None, ctx=None):
    """Field results for every frequency at once without storing M and L:
    (scattered, incident), each (nfreq, nfp); scattered holds the total field
    when predata.output_total, and both are pressures when
    predata.output_pressure (analysis.rs:100-111, postprocess.rs:70-88)."""
    ctx = ctx or capi.default_context(predata.device.index or 0)
    fs = predata.get_frequencies() if frequencies is None else list(frequencies)
    if predata.d_fps is None or predata.d_fps.shape[0] == 0:
        raise ValueError("no field points: call predata.set_field_points")
    n, nfp = predata.get_num_eqn(), predata.d_fps.shape[0]
    phi = phi.to(torch.complex128).reshape(len(fs), n).contiguous()
    vn = vn.to(torch.complex128).reshape(len(fs), n).contiguous()
    out = torch.empty((len(fs), nfp), dtype=torch.complex128, device=predata.device)
    inc = torch.empty((len(fs), nfp), dtype=torch.complex128, device=predata.device)
    w = predata.incident_wave.to_c()
    ctx.call("ccab_field", predata.physics_array(fs), len(fs), ct.byref(w), _ptr(predata.d_cpts),
             n, _ptr(predata.d_fps), nfp, _ptr(phi), _ptr(vn), int(predata.output_total),
             int(predata.output_pressure), _ptr(out), _ptr(inc))
    ctx.synchronize()
    return out, inc
