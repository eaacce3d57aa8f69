"""Bodies of uneven size solved in one call (ccab_solve_bodies): per-body
assembly and right-hand sides packed back to back, one GMRES-IR over all of
them with ragged batched matvecs.  Each body's solution is checked against
the oracle's FP64 LU of that body alone (solve.rs:11-71 run body by body), and
against the same solver run body by body."""
import math
import os

import numpy as np
import pytest
import torch

import oracle as O
from cc_abstemious_amd import capi, mesh, solve
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[case] {msg}", flush=True)


def _body(seed, nu):
    c = np.random.default_rng(seed).standard_normal(81) * 0.03
    c[0] = 0.0
    return mesh.perturbed_sphere(nu, c)


def _minus_one_element(m):
    """The mesh without its last element: an odd number of unknowns (an open
    surface, still a well-posed collocation system for the test)."""
    return mesh.Mesh(m.nodes, m.conn[:-1], m.nverts[:-1])


def _bodies(specs, bc=capi.BC_NORMAL_VELOCITY, bcv=(0.0, 0.0), solver=None, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for m in specs:
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        ka = rng.uniform(1.0, 6.0)
        f = ka * 343.0 / (2 * math.pi)
        out.append(PreData(m, [f], 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, bc, bcv,
                           IncidentWave(capi.WAVE_PLANE, tuple(d), (1.0, 0.0)),
                           solver or Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-12,
                                            restart=40, ir_tolerance=1e-3, ir_max_steps=20,
                                            ref_mgs=False, matrix_prec=capi.C32)))
    return out


def _oracle(pd):
    ph = O.physics(pd.get_frequency(), pd.sound_speed, pd.mass_density, pd.method,
                   pd.problem_type, pd.bc_type, pd.bc_value)
    w = O.wave(pd.incident_wave.kind, tuple(pd.incident_wave.vec),
               tuple(pd.incident_wave.amplitude))
    rhs = O.incident_surface(ph, w, pd.cpts)
    phi, vn, _ = O.solve_surface_dense(ph, pd.cpts, rhs)
    return phi, vn


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec,residual", [(capi.C32, capi.C128), (capi.C64, capi.C128),
                                            (capi.C32, capi.C32)])
def test_uneven_bodies_vs_oracle(prec, residual):
    """Five perturbed spheres of 320-1280 unknowns plus one of 1279 (odd: the
    one-entry-per-lane GEMV), rigid BC; every body to 1e-9 of its FP64 LU
    (1e-8 with residuals against the FP16 operator + FP32 remainder, whose
    entries carry about 2^-35 relative error; the FP16 operator alone would
    leave 1e-4-1e-3)."""
    specs = [_body(100 + k, nu) for k, nu in enumerate([4, 8, 5, 7, 6])]
    specs.append(_minus_one_element(mesh.icosphere(3)))
    solver = Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-12, restart=40,
                    ir_tolerance=1e-3, ir_max_steps=20, ref_mgs=False, matrix_prec=prec,
                    residual_matrix=residual)
    bodies = _bodies(specs, solver=solver, seed=1)
    bound = 1e-8 if residual == capi.C32 else 1e-9
    _note(f"bodies prec={prec} residual={residual} n={[pd.get_num_eqn() for pd in bodies]}")
    phis, vns, stats = solve.solve_bodies(bodies)
    for pd, phi, vn, st in zip(bodies, phis, vns, stats):
        assert st["converged"], st
        ref_phi, ref_vn = _oracle(pd)
        e = _rel(phi.cpu().numpy(), ref_phi)
        print(f"n={pd.get_num_eqn()} iterations {st['iterations']} ir {st['ir_steps']} err {e:.2e}")
        assert e < bound
        np.testing.assert_array_equal(vn.cpu().numpy(), ref_vn)   # rigid: vn = 0


def test_bodies_match_body_by_body_and_bcs():
    """Nonzero normal velocity and pressure BCs (their right-hand sides need
    the row sums of G or H): the batch equals the same solver body by body."""
    specs = [_body(200 + k, nu) for k, nu in enumerate([5, 7, 4])]
    for bc, bcv in [(capi.BC_NORMAL_VELOCITY, (1.0, 0.5)), (capi.BC_PRESSURE, (0.3, -0.2))]:
        bodies = _bodies(specs, bc=bc, bcv=bcv, seed=2)
        _note(f"bodies bc={bc} vs body by body")
        phis, vns, stats = solve.solve_bodies(bodies)
        for pd, phi, vn in zip(bodies, phis, vns):
            p1, v1, _ = solve.sweep(pd)
            assert _rel(phi.cpu().numpy(), p1[0].cpu().numpy()) < 1e-8
            assert _rel(vn.cpu().numpy(), v1[0].cpu().numpy()) < 1e-8
            ref_phi, ref_vn = _oracle(pd)
            assert _rel(phi.cpu().numpy(), ref_phi) < 1e-9
            assert _rel(vn.cpu().numpy(), ref_vn) < 1e-9


def test_bodies_direct_and_warm_start():
    """Body-by-body solvers (FP64 LU) through the same entry point, and a
    warm start: the FP64 LU answer perturbed by 1% starts GMRES-IR closer and
    it needs fewer iterations for the same answer."""
    specs = [_body(300 + k, nu) for k, nu in enumerate([6, 4, 7])]
    bodies = _bodies(specs, seed=3)
    _note("bodies direct + warm start")
    lu_phi, _, st_lu = solve.solve_bodies(bodies, solver=Solver(kind=capi.SOLVER_DIRECT))
    for pd, phi in zip(bodies, lu_phi):
        assert _rel(phi.cpu().numpy(), _oracle(pd)[0]) < 1e-11
    cold, _, st0 = solve.solve_bodies(bodies)
    rng = np.random.default_rng(4)
    x0 = [p * (1 + 0.01 * torch.from_numpy(rng.standard_normal(p.numel())).to(p.device))
          for p in lu_phi]
    warm, _, st1 = solve.solve_bodies(bodies, x0=x0)
    for a, b, s0, s1 in zip(cold, warm, st0, st1):
        assert s1["converged"]
        assert _rel(b.cpu().numpy(), a.cpu().numpy()) < 1e-8
        assert s1["iterations"] < s0["iterations"]


def test_bodies_in_memory_chunks():
    """A memory budget smaller than the batch: bodies go through GMRES-IR in
    several chunks, with the same answers."""
    specs = [_body(400 + k, nu) for k, nu in enumerate([6, 5, 7, 4, 6])]
    bodies = _bodies(specs, seed=5)
    _note("bodies in chunks (CCAB_MEM_BUDGET_MB=20)")
    full, _, _ = solve.solve_bodies(bodies)
    os.environ["CCAB_MEM_BUDGET_MB"] = "20"     # ~ two bodies of this size per chunk
    try:
        ctx = capi.Context()
    finally:
        del os.environ["CCAB_MEM_BUDGET_MB"]
    chunked, _, st = solve.solve_bodies(bodies, ctx=ctx)
    ctx.close()
    for a, b, s in zip(full, chunked, st):
        assert s["converged"]
        assert _rel(b.cpu().numpy(), a.cpu().numpy()) < 1e-8


@pytest.mark.parametrize("odd", [False, True])
def test_ragged_column_chunks(odd):
    """Bodies of 80-1280 unknowns with 64-column chunks of the ragged matvec
    (the small bodies' rows and chunks past their size write no partial
    sums) and with the workgroup-target split: every body to 1e-8 of its
    oracle FP64 LU either way.  odd: one body of 1279 unknowns (one entry per
    lane); else all sizes multiples of 4 (four rows per lane)."""
    specs = [_body(600, 2), _body(601, 8), _body(602, 4)]
    if odd:
        specs.append(_minus_one_element(mesh.icosphere(3)))
    solver = Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-12, restart=40,
                    ir_tolerance=1e-3, ir_max_steps=20, ref_mgs=False, matrix_prec=capi.C32,
                    residual_matrix=capi.C32)
    bodies = _bodies(specs, solver=solver, seed=7)
    refs = [_oracle(pd)[0] for pd in bodies]
    out = {}
    for chunk in ("64", "0"):
        os.environ["CCAB_GEMV_RAGGED_CHUNK"] = chunk
        try:
            ctx = capi.Context()
        finally:
            del os.environ["CCAB_GEMV_RAGGED_CHUNK"]
        _note(f"ragged chunk={chunk} odd={odd} n={[pd.get_num_eqn() for pd in bodies]}")
        phis, _, st = solve.solve_bodies(bodies, ctx=ctx)
        out[chunk] = [p.cpu().numpy() for p in phis]
        ctx.close()
        for pd, phi, s_, ref in zip(bodies, out[chunk], st, refs):
            assert s_["converged"], s_
            assert _rel(phi, ref) < 1e-8, (pd.get_num_eqn(), _rel(phi, ref))
    for a, b in zip(out["64"], out["0"]):
        assert _rel(a, b) < 1e-8


def test_bodies_errors():
    specs = [_body(500, 4), _body(501, 5)]
    bodies = _bodies(specs, seed=6)
    bodies[1].bc_type = capi.BC_PRESSURE
    with pytest.raises(capi.CcabError):
        solve.solve_bodies(bodies)
    bodies = _bodies(specs, seed=6)
    with pytest.raises(ValueError):
        solve.solve_bodies(bodies, x0=[torch.zeros(3, dtype=torch.complex128)] * 2)
