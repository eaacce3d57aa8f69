"""A warm start's first residual b - A x0 with A x0 formed inside the
assembly (each FP64 entry times the guess as the kernel forms it; no stored
matrix is read) against A x0 as a product over the stored matrices after the
assembly (ccab_set_warm_ax): the same solutions to the solver's tolerance,
the same iteration counts to within one Arnoldi step per system, and both
within 1e-8 of the FP64 LU."""
import math
import sys

import numpy as np
import pytest
import torch

from cc_abstemious_amd import capi, mesh, solve
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[warm-ax] {msg}", file=sys.stderr, flush=True)


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("bc,bcv,residual", [
    (capi.BC_NORMAL_VELOCITY, (0.0, 0.0), capi.C32),    # H-only column loop, FP16 + FP32 remainder
    (capi.BC_NORMAL_VELOCITY, (0.0, 0.0), capi.C128),
    (capi.BC_IMPEDANCE, (415.03, 83.006), capi.C32),    # fused lhs_a H + lhs_b G loop
    (capi.BC_NORMAL_VELOCITY, (1.0, 0.5), capi.C32),    # row sums and A x0 in LDS together
])
def test_sweep_warm_ax_matches_product(bc, bcv, residual):
    _note(f"sweep warm start, A x0 from the assembly vs product; bc={bc} bcv={bcv} residual={residual}")
    m = mesh.icosphere(3)
    ka = np.linspace(0.5, 8.0, 20)       # 20 frequencies: a 16-system group and a partial one
    fs = ka * 343.0 / (2 * math.pi)
    pd = PreData(m, fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, bc, bcv,
                 IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_DIRECT))
    phi_d, _, _ = solve.sweep(pd)
    g = torch.Generator().manual_seed(21)
    x0 = phi_d * (1 + 0.01 * torch.randn(phi_d.shape, generator=g,
                                         dtype=torch.float64).to(phi_d.device))
    pd.solver = Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=30,
                       ir_tolerance=1e-3, ir_max_steps=30, ref_mgs=False, matrix_prec=capi.C32,
                       residual_matrix=residual)
    ctx = capi.Context()
    try:
        phi_a, _, st_a = solve.sweep(pd, x0=x0, ctx=ctx)
        ctx.set_warm_ax(False)
        phi_p, _, st_p = solve.sweep(pd, x0=x0, ctx=ctx)
        ctx.set_warm_ax(True)
        phi_a2, _, _ = solve.sweep(pd, x0=x0, ctx=ctx)
    finally:
        ctx.close()
    assert torch.equal(phi_a, phi_a2)     # deterministic
    for k in range(len(fs)):
        assert st_a[k]["converged"] == 1 and st_a[k]["rel_residual"] <= 1e-10
        assert _rel(phi_a[k], phi_d[k]) < 1e-8, (k, _rel(phi_a[k], phi_d[k]))
        assert _rel(phi_a[k], phi_p[k]) < 1e-9
        assert abs(st_a[k]["iterations"] - st_p[k]["iterations"]) <= 1, (k, st_a[k], st_p[k])
    print("iterations", sum(s["iterations"] for s in st_a), sum(s["iterations"] for s in st_p))


def test_bodies_warm_ax_matches_product():
    _note("bodies warm start, A x0 from the assembly vs product")
    rng = np.random.default_rng(8)
    bodies = []
    for k, lvl in enumerate([2, 3, 2]):
        m = mesh.icosphere(lvl)
        if k == 2:   # an odd size
            m = mesh.Mesh(m.nodes, m.conn[:-1], m.nverts[:-1])
        d = rng.standard_normal(3)
        f = (1.0 + 5.0 * rng.random()) * 343.0 / (2 * math.pi)
        bodies.append(PreData(m, [f], 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR,
                              capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                              IncidentWave(capi.WAVE_PLANE, tuple(d / np.linalg.norm(d)), (1.0, 0.0)),
                              Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=30,
                                     ir_tolerance=1e-3, ir_max_steps=30, ref_mgs=False,
                                     matrix_prec=capi.C32, residual_matrix=capi.C32)))
    lu, _, _ = solve.solve_bodies(bodies, solver=Solver(kind=capi.SOLVER_DIRECT))
    x0 = [p * (1 + 0.01 * torch.from_numpy(rng.standard_normal(p.numel())).to(p.device)) for p in lu]
    ctx = capi.Context()
    try:
        wa, _, sa = solve.solve_bodies(bodies, x0=x0, ctx=ctx)
        ctx.set_warm_ax(False)
        wp, _, sp = solve.solve_bodies(bodies, x0=x0, ctx=ctx)
    finally:
        ctx.close()
    for a, p, ref, s_a, s_p in zip(wa, wp, lu, sa, sp):
        assert s_a["converged"]
        assert _rel(a, ref) < 1e-8
        assert _rel(a, p) < 1e-9
        assert abs(s_a["iterations"] - s_p["iterations"]) <= 1
