"""ctypes binding of libccab.so (the C ABI declared in include/ccab.h).

The library is the product: importing this module raises if it has not been
built (``python -m cc_abstemious_amd.build``) -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccab.so")

OK, EINVAL, EHIP, ENOMEM, ENAN, ESINGULAR = 0, -1, -2, -3, -4, -5
CLASSICAL, BURTON_MILLER = 0, 1
EXTERIOR, INTERIOR = 0, 1
BC_PRESSURE, BC_NORMAL_VELOCITY, BC_IMPEDANCE = 0, 1, 2
WAVE_PLANE, WAVE_SPHERICAL = 0, 1
LHS, RHS = 0, 1
C128, C64, C32 = 0, 1, 2
SOLVER_DIRECT, SOLVER_ITERATIVE, SOLVER_HIERARCHICAL = 0, 1, 2
SOLVER_DIRECT_MIXED, SOLVER_ITERATIVE_MIXED = 3, 4
STOP_BACKWARD_ERROR, STOP_RELATIVE_RESIDUAL = 0, 1
# physics corrections (opt-in; 0 = the reference's numerics), include/ccab.h
FIX_BM_DH, FIX_FIELD_SIGN, FIX_IMPEDANCE_VN, FIX_SPHERICAL_ORIGIN, FIX_QUAD_AREA = 1, 2, 4, 8, 16
FIX_PRESSURE_SIGN = 32
FIX_ALL = 63
# LU GEMM variants (ccab_set_tuning): 0..7 the MFMA kernel, 8 rocBLAS
GEMM_VARIANTS, GEMM_LIBRARY = 9, 8


class CcabError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ccab error {code}: {msg}")
        self.code = code


class Physics(ct.Structure):
    _fields_ = [
        ("frequency", ct.c_double),
        ("sound_speed", ct.c_double),
        ("mass_density", ct.c_double),
        ("method", ct.c_int32),
        ("problem_type", ct.c_int32),
        ("bc_type", ct.c_int32),
        ("corrections", ct.c_int32),
        ("bc_value", ct.c_double * 2),
    ]


class Wave(ct.Structure):
    _fields_ = [
        ("kind", ct.c_int32),
        ("pad_", ct.c_int32),
        ("vec", ct.c_double * 3),
        ("amplitude", ct.c_double * 2),
    ]


class GmresParams(ct.Structure):
    _fields_ = [
        ("max_iterations", ct.c_int64),
        ("tolerance", ct.c_double),
        ("restart", ct.c_int32),
        ("ref_mgs", ct.c_int32),
        ("stop_rule", ct.c_int32),
        ("matrix_prec", ct.c_int32),
    ]


class SolveStats(ct.Structure):
    _fields_ = [
        ("num_mv", ct.c_int64),
        ("iterations", ct.c_int64),
        ("restarts", ct.c_int64),
        ("ir_steps", ct.c_int64),
        ("backward_error", ct.c_double),
        ("rel_residual", ct.c_double),
        ("converged", ct.c_int32),
        ("pad_", ct.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad_"}


class SolverCfg(ct.Structure):
    _fields_ = [
        ("solver", ct.c_int32),
        ("leaf_size", ct.c_int32),
        ("gmres", GmresParams),
        ("ir_tolerance", ct.c_double),
        ("ir_max_steps", ct.c_int32),
        ("residual_matrix", ct.c_int32),
        ("aca_tolerance", ct.c_double),
        ("eta", ct.c_double),
    ]


class SurrogateCfg(ct.Structure):
    _fields_ = [
        ("width", ct.c_int32),
        ("seed", ct.c_int32),
        ("lr", ct.c_double),
        ("beta1", ct.c_double),
        ("beta2", ct.c_double),
        ("eps", ct.c_double),
        ("weight_decay", ct.c_double),
    ]


# name -> (restype, argtypes); every exported symbol of include/ccab.h
_P = ct.c_void_p
_I64 = ct.c_int64
_I = ct.c_int
SIGNATURES = {
    "ccab_version": (_I, []),
    "ccab_create": (_I, [_I, ct.POINTER(_P)]),
    "ccab_destroy": (_I, [_P]),
    "ccab_last_error": (ct.c_char_p, [_P]),
    "ccab_set_stream": (_I, [_P, _P]),
    "ccab_synchronize": (_I, [_P]),
    "ccab_profile_enable": (_I, [_P, _I]),
    "ccab_profile_read": (_I, [_P, _I, _P, _P, _P, _I]),
    "ccab_set_tuning": (_I, [_P, _I, _I, _I]),
    "ccab_set_lu_coop": (_I, [_P, _I64]),
    "ccab_stream_join": (_I, [_P, _P]),
    "ccab_autotune": (_I, [_P, _I, _I64, _I64, _P, _P, _P]),
    "ccab_set_graphs": (_I, [_P, _I]),
    "ccab_set_warm_ax": (_I, [_P, _I]),
    "ccab_set_gir_deflation": (_I, [_P, _I]),
    "ccab_stream_wait": (_I, [_P, _P]),
    "ccab_collocation_points": (_I, [_P, _P, _P, _I64, _P]),
    "ccab_collocation_points_ex": (_I, [_P, _P, _P, _I64, _I, _P]),
    "ccab_incident_surface": (_I, [_P, ct.POINTER(Physics), ct.POINTER(Wave), _P, _I64, _P]),
    "ccab_dense_surface_matrices": (_I, [_P, ct.POINTER(Physics), _P, _I64, _I, _P, _P, _I64]),
    "ccab_assemble_system": (_I, [_P, _P, _I64, _P, _I64, _P, _P, _I64, _I64, _P, _P]),
    "ccab_surface_row_or_column": (_I, [_P, ct.POINTER(Physics), _P, _I64, _I, _P, _I64, _P,
                                        _I64, _P]),
    "ccab_solve_lu": (_I, [_P, _I64, _P, _I64, _P]),
    "ccab_lu_factor_batched": (_I, [_P, _I, _I64, _I64, _P, _I64, _I64, _P]),
    "ccab_lu_solve_batched": (_I, [_P, _I, _I64, _I64, _P, _I64, _I64, _P, _P, _I64, _I64,
                                   _I64]),
    "ccab_solve_mixed_batched": (_I, [_P, _I64, _I64, _P, _P, _I64, _I64, _P, ct.c_double, _I,
                                      _P]),
    "ccab_gmres_dense": (_I, [_P, _I64, _P, _I64, _P, _P, ct.POINTER(GmresParams), _P]),
    "ccab_solve_surface": (_I, [_P, ct.POINTER(Physics), _P, _I64, _P, ct.POINTER(SolverCfg),
                                _P, _P, _P]),
    "ccab_sweep": (_I, [_P, _P, _I64, ct.POINTER(Wave), _P, _I64, ct.POINTER(SolverCfg), _P, _P,
                        _P]),
    "ccab_cgemm_update": (_I, [_P, _I, _I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64]),
    "ccab_gmres_ir_batched": (_I, [_P, _I64, _I64, _P, _P, _I64, _I64, _P, _P,
                                   ct.POINTER(SolverCfg), _P]),
    "ccab_sweep_warm": (_I, [_P, _P, _I64, ct.POINTER(Wave), _P, _I64, ct.POINTER(SolverCfg), _P,
                             _P, _P, _P]),
    "ccab_solve_bodies": (_I, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ccab_hmatrix_build": (_I, [_P, ct.POINTER(Physics), _P, _I64, _I, _I64, ct.c_double,
                                ct.c_double, ct.POINTER(_P)]),
    "ccab_hmatrix_build_ex": (_I, [_P, ct.POINTER(Physics), _P, _I64, _I, _I64, ct.c_double,
                                   ct.c_double, _I, ct.POINTER(_P)]),
    "ccab_hmatrix_destroy": (_I, [_P]),
    "ccab_hmatrix_gemv": (_I, [_P, _P, _P, _P, _P, _P]),
    "ccab_hmatrix_to_full": (_I, [_P, _P, _P, _I64]),
    "ccab_hmatrix_stats": (_I, [_P, _P, _P]),
    "ccab_hmatrix_stats_ex": (_I, [_P, _P, _I, _P]),
    "ccab_hmatrix_build_matrix": (_I, [_P, _I64, _P, _I64, _P, _I64, ct.c_double, ct.c_double,
                                       _I, ct.POINTER(_P)]),
    "ccab_aca_build": (_I, [_P, _I64, _I64, _P, _I64, ct.c_double, ct.POINTER(_P)]),
    "ccab_gmres_hmatrix": (_I, [_P, _P, _P, _P, ct.POINTER(GmresParams), _P]),
    "ccab_field": (_I, [_P, _P, _I64, _P, _P, _I64, _P, _I64, _P, _P, _I, _I, _P, _P]),
    "ccab_autotune_sweep": (_I, [_P, _P, _I64, _P, _I64, _I, _P, _P]),
    "ccab_set_sweep_tuning": (_I, [_P, _I, _I64, _P]),
    "ccab_field_noalloc": (_I, [_P, _P, _I64, _P, _P, _I64, _P, _I64, _P, _P, _I, _I, _P, _P]),
    "ccab_dense_field_matrices": (_I, [_P, ct.POINTER(Physics), _P, _I64, _P, _I64, _P, _P, _I64]),
    "ccab_incident_field": (_I, [_P, ct.POINTER(Physics), ct.POINTER(Wave), _P, _I64, _P]),
    "ccab_surrogate_create": (_I, [_P, ct.POINTER(SurrogateCfg), ct.POINTER(_P)]),
    "ccab_surrogate_destroy": (_I, [_P]),
    "ccab_surrogate_num_params": (_I64, [_P]),
    "ccab_surrogate_get_params": (_I, [_P, _P, _P]),
    "ccab_surrogate_set_params": (_I, [_P, _P, _P]),
    "ccab_surrogate_set_lr": (_I, [_P, ct.c_double]),
    "ccab_surrogate_features": (_I, [_P, _P, _P, _I64, _P, _I64, _P, _P, _P]),
    "ccab_surrogate_loss_grad": (_I, [_P, _P, _P, _P, _I64, _P, _P]),
    "ccab_surrogate_train_step": (_I, [_P, _P, _P, _P, _P, _I64, _P]),
    "ccab_surrogate_forward": (_I, [_P, _P, _P, _I64, _P]),
    "ccab_surrogate_predict": (_I, [_P, _P, _P, _P, _I64, _P, _I64, _P]),
    "ccab_surrogate_predict_bodies": (_I, [_P, _P, _I64, _P, _P, _P, _P, _P]),
}

_lib = None


def lib():
    """Load libccab.so; raises ImportError if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cc_abstemious_amd.build` "
                "(there is no CPU fallback)")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def physics(frequency, sound_speed=343.0, mass_density=1.21, method=CLASSICAL,
            problem_type=EXTERIOR, bc_type=BC_NORMAL_VELOCITY, bc_value=(0.0, 0.0),
            corrections=0) -> Physics:
    ph = Physics()
    ph.corrections = int(corrections)
    ph.frequency = float(frequency)
    ph.sound_speed = float(sound_speed)
    ph.mass_density = float(mass_density)
    ph.method = int(method)
    ph.problem_type = int(problem_type)
    ph.bc_type = int(bc_type)
    ph.bc_value[0], ph.bc_value[1] = (float(bc_value[0]), float(bc_value[1]))
    return ph


def wave(kind=WAVE_PLANE, vec=(1.0, 0.0, 0.0), amplitude=(1.0, 0.0)) -> Wave:
    w = Wave()
    w.kind = int(kind)
    for d in range(3):
        w.vec[d] = float(vec[d])
    w.amplitude[0], w.amplitude[1] = (float(amplitude[0]), float(amplitude[1]))
    return w


def gmres_params(max_iterations=1000, tolerance=1e-5, restart=0, ref_mgs=True,
                 stop_rule=STOP_BACKWARD_ERROR, matrix_prec=C128) -> GmresParams:
    p = GmresParams()
    p.max_iterations = int(max_iterations)
    p.tolerance = float(tolerance)
    p.restart = int(restart)
    p.ref_mgs = int(bool(ref_mgs))
    p.stop_rule = int(stop_rule)
    p.matrix_prec = int(matrix_prec)
    return p


def solver_cfg(solver=SOLVER_DIRECT, gmres: GmresParams | None = None, ir_tolerance=1e-12,
               ir_max_steps=10, leaf_size=32, aca_tolerance=1e-4, eta=4.0,
               residual_matrix=C128) -> SolverCfg:
    c = SolverCfg()
    c.solver = int(solver)
    c.gmres = gmres if gmres is not None else gmres_params()
    c.ir_tolerance = float(ir_tolerance)
    c.ir_max_steps = int(ir_max_steps)
    c.residual_matrix = int(residual_matrix)
    c.leaf_size = int(leaf_size)
    c.aca_tolerance = float(aca_tolerance)
    c.eta = float(eta)
    return c


def _torch_stream():
    """HIP stream handle of torch's current stream (None = legacy default)."""
    import torch
    if not torch.cuda.is_initialized():
        return None
    h = torch.cuda.current_stream().cuda_stream
    return ct.c_void_p(h) if h else None


class Context:
    """One device context (HIP stream + workspace) of libccab."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = ct.c_void_p()
        rc = self._L.ccab_create(device, ct.byref(h))
        if rc != OK:
            raise CcabError(rc, f"ccab_create(device={device}) failed")
        self.handle = h
        self.device = device

    def check(self, rc: int):
        if rc != OK:
            raise CcabError(rc, self._L.ccab_last_error(self.handle).decode())

    def call(self, name: str, *args):
        # inputs produced by torch are queued on torch's current stream: order
        # the library's stream after them
        ts = _torch_stream()
        self.check(self._L.ccab_stream_wait(self.handle, ts))
        self.check(getattr(self._L, name)(self.handle, *args))
        # and torch's stream after the library's: results are read there next
        self.check(self._L.ccab_stream_join(self.handle, ts))

    def autotune(self, prec: int, n: int, batch: int = 1):
        """Pick the fastest LU GEMM variant and outer panel width by timing;
        returns (variant, outer, {(outer, variant): ms})."""
        v, o = ct.c_int(), ct.c_int()
        nv = GEMM_VARIANTS
        tab = (ct.c_double * (2 * nv))()
        self.call("ccab_autotune", int(prec), int(n), int(batch), ct.byref(v), ct.byref(o), tab)
        times = {(64 if i < nv else 128, i % nv): tab[i] for i in range(2 * nv)}
        return v.value, o.value, times

    def autotune_sweep(self, predata, op_prec: int = C32, frequencies=None):
        """Tune the batched assembly and GEMVs of predata's sweep (its mesh and
        frequencies); returns (best settings dict, timing table dict)."""
        fs = predata.get_frequencies() if frequencies is None else list(frequencies)
        best = (ct.c_int * 6)()
        tab = (ct.c_double * 29)()
        self.call("ccab_autotune_sweep", predata.physics_array(fs), len(fs),
                  ct.c_void_p(predata.d_cpts.data_ptr()), predata.get_num_eqn(), int(op_prec),
                  best, tab)
        keys = ["systems_per_thread", "column_chunk", "op_wg_target", "op_nt", "ref_wg_target",
                "ref_nt"]
        t = list(tab)
        table = {"assembly_ms": {(kf, cc): t[a * 3 + c] for a, kf in enumerate((8, 16, 32))
                                 for c, cc in enumerate((64, 128, 256))},
                 "op_gemv_ms": {(wg, nt): t[9 + w * 2 + nt] for w, wg in
                                enumerate((256, 512, 1024, 2048, 4096)) for nt in (0, 1)},
                 "ref_gemv_ms": {(wg, nt): t[19 + w * 2 + nt] for w, wg in
                                 enumerate((256, 512, 1024, 2048, 4096)) for nt in (0, 1)}}
        return dict(zip(keys, list(best))), table

    def set_sweep_tuning(self, op_prec: int, n: int, cfg=None):
        """Force the sweep tuning of size n (6 ints as autotune_sweep's best, in
        its key order) or drop it (None)."""
        arr = (ct.c_int * 6)(*cfg) if cfg is not None else None
        self.call("ccab_set_sweep_tuning", int(op_prec), int(n), arr)

    def set_graphs(self, on: bool):
        """Replay captured HIP graphs of the LU factorisation."""
        self.call("ccab_set_graphs", int(bool(on)))

    def set_warm_ax(self, on: bool):
        """Form a warm start's A x0 inside the assembly (default) or as a
        product over the stored matrices."""
        self.call("ccab_set_warm_ax", int(bool(on)))

    def set_gir_deflation(self, nvec: int):
        """Refinement cycles of the batched GMRES-IR after each system's first
        deflated by nvec directions of its first cycle's Krylov space (default
        8, harmonic Ritz directions); 0: every cycle restarts from scratch."""
        self.call("ccab_set_gir_deflation", int(nvec))

    def set_tuning(self, prec: int, variant: int, outer: int):
        self.call("ccab_set_tuning", int(prec), int(variant), int(outer))

    def set_lu_coop(self, min_rows: int):
        """Panels of at least min_rows rows use the multi-workgroup panel
        factorisation when the batch is small (0: never)."""
        self.call("ccab_set_lu_coop", int(min_rows))

    def synchronize(self):
        self.check(self._L.ccab_synchronize(self.handle))

    def set_stream(self, stream_ptr: int | None):
        self.check(self._L.ccab_set_stream(self.handle, ct.c_void_p(stream_ptr or 0)))

    PROFILE_CLASSES = ["assembly", "lu_panel", "lu_trsm", "lu_update", "lu_laswp", "lu_trinv",
                       "trsv", "gemv", "mgs", "field", "hmatrix_gemv", "surrogate", "gemv_c128"]

    def profile(self, on: bool = True):
        self.check(self._L.ccab_profile_enable(self.handle, int(on)))

    def profile_read(self, reset: bool = True) -> dict:
        out = {}
        for k, name in enumerate(self.PROFILE_CLASSES):
            cnt = ct.c_int64()
            ms = ct.c_double()
            work = ct.c_double()
            self.check(self._L.ccab_profile_read(self.handle, k, ct.byref(cnt), ct.byref(ms),
                                                 ct.byref(work), int(reset)))
            if cnt.value:
                out[name] = {"count": cnt.value, "ms": ms.value, "work": work.value}
        return out

    def close(self):
        if getattr(self, "handle", None):
            self._L.ccab_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]
