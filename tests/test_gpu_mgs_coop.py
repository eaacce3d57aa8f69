"""The Gram-Schmidt step of GMRES for few large systems (mgs_coop_kernel: one
system's orthogonalisation over many workgroups, a grid barrier per basis
vector) against the one-workgroup kernel (environment CCAB_MGS_COOP=0): the
same iterations and matvec counts and the same solution to rounding, for the
reference's GMRES (coefficient of gmres.rs:168) and the textbook variant, the
batched GMRES-IR and the H-matrix GMRES."""
import math
import os

import numpy as np
import pytest
import torch

from cc_abstemious_amd import capi, mesh, solve
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx_pair():
    os.environ["CCAB_MGS_COOP"] = "0"
    try:
        one = capi.Context()
    finally:
        del os.environ["CCAB_MGS_COOP"]
    coop = capi.Context()
    yield one, coop
    one.close()
    coop.close()


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("ref_mgs", [True, False])
def test_dense_gmres(ctx_pair, ref_mgs):
    """A BEM system of 5120 unknowns (Burton-Miller, ka = 4) with the
    reference's restart and stop rule."""
    one, coop = ctx_pair
    pd = PreData(mesh.icosphere(4), [4.0 * 343.0 / (2 * math.pi)], 343.0, 1.21,
                 capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                 IncidentWave(capi.WAVE_PLANE, (0.0, 0.6, 0.8), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_ITERATIVE, tolerance=1e-10, max_iterations=1000,
                        ref_mgs=ref_mgs))
    _note(f"GMRES n=5120 ref_mgs={ref_mgs}: cooperative vs one-workgroup Gram-Schmidt")
    out = []
    for ctx in (one, coop):
        phi, _, st = solve.sweep(pd, ctx=ctx)
        out.append((phi[0], st[0]))
    (p1, s1), (p2, s2) = out
    print(s1, s2)
    assert s1["iterations"] == s2["iterations"] and s1["num_mv"] == s2["num_mv"]
    assert s1["converged"] and s2["converged"]
    assert _rel(p2, p1) < 1e-9


def test_gmres_ir_few_systems(ctx_pair):
    """Three frequencies of the sweep (3 x 20 workgroups per Gram-Schmidt)."""
    one, coop = ctx_pair
    ka = np.array([1.0, 3.5, 7.0])
    pd = PreData(mesh.icosphere(4), ka * 343.0 / (2 * math.pi), 343.0, 1.21,
                 capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                 IncidentWave(capi.WAVE_PLANE, (0.3, -0.5, 0.8), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=30,
                        ir_tolerance=1e-3, ir_max_steps=30, ref_mgs=False,
                        matrix_prec=capi.C32))
    _note("GMRES-IR 3 systems n=5120: cooperative vs one-workgroup Gram-Schmidt")
    r1 = solve.sweep(pd, ctx=one)
    r2 = solve.sweep(pd, ctx=coop)
    for a, b, sa, sb in zip(r1[0], r2[0], r1[2], r2[2]):
        assert sa["converged"] and sb["converged"]
        assert abs(sa["iterations"] - sb["iterations"]) <= 1
        assert _rel(b, a) < 1e-8


def test_hmatrix_gmres(ctx_pair):
    """Hierarchical solve (n = 5120, leaf 64, eta 1): the GMRES on the H-matrix
    operator with the reference's stop rule."""
    one, coop = ctx_pair
    pd = PreData(mesh.icosphere(4), [6.0 * 343.0 / (2 * math.pi)], 343.0, 1.21,
                 capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                 IncidentWave(capi.WAVE_PLANE, (1.0, 0.0, 0.0), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_HIERARCHICAL, tolerance=1e-6, max_iterations=1000,
                        leaf_size=64, eta=1.0))
    _note("H-matrix GMRES n=5120: cooperative vs one-workgroup Gram-Schmidt")
    r1 = solve.sweep(pd, ctx=one)
    r2 = solve.sweep(pd, ctx=coop)
    assert r1[2][0]["num_mv"] == r2[2][0]["num_mv"]
    assert _rel(r2[0][0], r1[0][0]) < 1e-9


@pytest.mark.parametrize("coop_on", [True, False])
def test_gmres_ir_blocked_gram_schmidt(coop_on):
    """The batched GMRES-IR's Gram-Schmidt with the basis vectors 8 (the
    default) or 4 at a time per reduction (classical within a block, modified
    across blocks) against one at a time (CCAB_GS_BLOCK=1), on the
    multi-workgroup path (3 systems of 5120 unknowns) and on one workgroup
    per system (CCAB_MGS_COOP=0): the same iterations to within one per
    system and the same solutions to the solver's tolerance."""
    ka = np.array([1.0, 3.5, 7.0])
    pd = PreData(mesh.icosphere(4), ka * 343.0 / (2 * math.pi), 343.0, 1.21,
                 capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                 IncidentWave(capi.WAVE_PLANE, (0.3, -0.5, 0.8), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=30,
                        ir_tolerance=1e-3, ir_max_steps=30, ref_mgs=False,
                        matrix_prec=capi.C32))
    _note(f"GMRES-IR blocked Gram-Schmidt, multi-workgroup={coop_on}")
    res = {}
    for blk in ("1", "4", "8"):
        os.environ["CCAB_GS_BLOCK"] = blk
        if not coop_on:
            os.environ["CCAB_MGS_COOP"] = "0"
        try:
            ctx = capi.Context()
        finally:
            del os.environ["CCAB_GS_BLOCK"]
            os.environ.pop("CCAB_MGS_COOP", None)
        res[blk] = solve.sweep(pd, ctx=ctx)
        ctx.close()
    for blk in ("4", "8"):
        for a, b, sa, sb in zip(res["1"][0], res[blk][0], res["1"][2], res[blk][2]):
            assert sa["converged"] and sb["converged"]
            assert sb["rel_residual"] <= 1e-10
            assert abs(sa["iterations"] - sb["iterations"]) <= 1, (blk, sa, sb)
            assert _rel(b, a) < 1e-8
