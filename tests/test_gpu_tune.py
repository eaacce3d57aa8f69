"""The LU autotuner: it returns the fastest measured configuration, and the
factors do not depend on the GEMM variant it picks (results invariant to
tile choice); the outer panel width changes only the rounding."""
import numpy as np
import pytest
import torch

from cc_abstemious_amd import capi, solve

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("prec", [capi.C64, capi.C128])
def test_autotune_picks_fastest(prec):
    _note(f"autotune prec={prec} n=1536 batch=2")
    ctx = capi.Context()
    v, o, times = ctx.autotune(prec, 1536, 2)
    assert 0 <= v < capi.GEMM_VARIANTS and o in (64, 128)
    assert len(times) == 2 * capi.GEMM_VARIANTS
    assert all(np.isfinite(t) and t > 0 for t in times.values())
    assert times[(o, v)] == min(times.values())
    print({k: round(t, 3) for k, t in times.items()})
    ctx.close()


@pytest.mark.parametrize("prec", [capi.C64, capi.C128])
def test_factors_invariant_to_gemm_variant(prec):
    n, batch = 700, 2
    _note(f"LU factors vs GEMM variant prec={prec} n={n}")
    dt = torch.complex64 if prec == capi.C64 else torch.complex128
    g = torch.Generator(device="cpu").manual_seed(5)
    A0 = (torch.randn(batch, n, n, dtype=torch.complex128, generator=g)
          + n ** 0.5 * torch.eye(n, dtype=torch.complex128)).to(dt).cuda()
    ctx = capi.Context()
    ref = None
    for outer in (128, 64):
        for v in range(8):
            ctx.set_tuning(prec, v, outer)
            A = A0.clone().transpose(1, 2).contiguous().transpose(1, 2)  # column-major per system
            piv = solve.lu_factor_batched(A, ctx=ctx)
            LU = A
            if ref is None:
                ref = (LU.clone(), piv.clone())
                continue
            assert torch.equal(piv, ref[1]), (outer, v)
            if outer == 128:
                assert torch.equal(LU, ref[0]), (outer, v)
            else:
                e = float((LU - ref[0]).abs().max() / ref[0].abs().max())
                assert e < (1e-5 if prec == capi.C64 else 1e-13), (outer, v, e)
    ctx.close()


@pytest.mark.parametrize("prec", [capi.C64, capi.C128])
def test_library_variant_matches_kernel(prec):
    """GEMM variant 8 (the trailing update by rocBLAS): the factorisation
    solves the systems as the MFMA kernel's does, to rounding (the pivots may
    differ only where a column holds near-ties)."""
    n, batch = 1000, 3
    _note(f"LU library variant vs MFMA kernel prec={prec} n={n} batch={batch}")
    dt = torch.complex64 if prec == capi.C64 else torch.complex128
    g = torch.Generator(device="cpu").manual_seed(7)
    A64 = torch.randn(batch, n, n, dtype=torch.complex128, generator=g) + \
        n ** 0.5 * torch.eye(n, dtype=torch.complex128)
    b64 = torch.randn(batch, n, dtype=torch.complex128, generator=g)
    ctx = capi.Context()
    xs = {}
    for v in (3, capi.GEMM_LIBRARY):
        for outer in (64, 128):
            ctx.set_tuning(prec, v, outer)
            A = A64.to(dt).cuda().transpose(1, 2).contiguous().transpose(1, 2)
            piv = solve.lu_factor_batched(A, ctx=ctx)
            x = solve.lu_solve_batched(A, piv, b64.to(dt).cuda(), ctx=ctx)
            xs[(v, outer)] = x.to(torch.complex128).cpu()
    eps = 1e-4 if prec == capi.C64 else 1e-12
    for key, x in xs.items():
        r = torch.linalg.vector_norm(A64 @ x.unsqueeze(-1) - b64.unsqueeze(-1)) / \
            torch.linalg.vector_norm(b64)
        d = torch.linalg.vector_norm(x - xs[(3, 128)]) / torch.linalg.vector_norm(xs[(3, 128)])
        print(key, float(r), float(d))
        assert float(r) < eps and float(d) < eps, key
    ctx.close()


@pytest.mark.parametrize("variant", [None, capi.GEMM_LIBRARY])
@pytest.mark.parametrize("prec", [capi.C64, capi.C128])
def test_lu_graph_replay_matches_eager(prec, variant):
    """Captured-and-replayed LU (HIP graph) equals the eager launches, over
    several replays with eager work and synchronisations in between.  With
    the library variant selected the graph holds the MFMA kernel's variant 3
    (rocBLAS is not captured), so it equals the eager variant-3 factors."""
    n, batch = 700, 3
    _note(f"LU graph replay prec={prec} n={n} batch={batch} variant={variant}")
    dt = torch.complex64 if prec == capi.C64 else torch.complex128
    g = torch.Generator(device="cpu").manual_seed(11)
    A0 = (torch.randn(batch, n, n, dtype=torch.complex128, generator=g)
          + n ** 0.5 * torch.eye(n, dtype=torch.complex128)).to(dt).cuda()
    ctx = capi.Context()
    if variant is not None:
        ctx.set_tuning(prec, 3, 128)
    work = A0.clone().transpose(1, 2).contiguous().transpose(1, 2)
    piv_e = solve.lu_factor_batched(work, ctx=ctx)
    LU_e = work.clone()
    if variant is not None:
        ctx.set_tuning(prec, variant, 128)
    ctx.set_graphs(True)
    for rep in range(4):
        work.copy_(A0.transpose(1, 2).contiguous().transpose(1, 2))
        piv = solve.lu_factor_batched(work, ctx=ctx)      # rep 0 captures, then replays
        assert torch.equal(piv, piv_e), rep
        assert torch.equal(work, LU_e), rep
        # unrelated eager work between replays
        b = torch.randn(batch, n, dtype=torch.complex128, generator=g).cuda().to(dt)
        solve.lu_solve_batched(work, piv, b.clone(), ctx=ctx)
        torch.cuda.synchronize()
    ctx.set_graphs(False)
    ctx.close()


def _sweep_case(level=3, nfreq=16, bc=capi.BC_NORMAL_VELOCITY, bcv=(0.0, 0.0)):
    import math
    from cc_abstemious_amd import mesh
    from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver
    fs = list(np.linspace(0.5, 8.0, nfreq) * 343.0 / (2 * math.pi))
    return PreData(mesh.icosphere(level), fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, bc,
                   bcv, IncidentWave(capi.WAVE_PLANE, (0.3, -0.4, 0.866), (1.0, 0.0)),
                   Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=30,
                          ir_tolerance=1e-3, ir_max_steps=60, ref_mgs=False,
                          stop_rule=capi.STOP_RELATIVE_RESIDUAL, matrix_prec=capi.C32,
                          residual_matrix=capi.C32))


def test_sweep_autotune_and_results_invariant():
    """ccab_autotune_sweep keeps the fastest measured settings; the sweep's
    solutions with the tuned and with extreme hand-set settings agree with
    the untuned ones to the solve tolerance (the settings only change how
    the work is split)."""
    pd = _sweep_case()
    n = pd.get_num_eqn()
    _note(f"sweep autotune n={n} nfreq=16 (FP16 operator, FP32 remainder)")
    ctx = capi.Context()
    phi0, _, st0 = solve.sweep(pd, ctx=ctx)
    best, table = ctx.autotune_sweep(pd, capi.C32)
    print(best)
    for key, t in table.items():
        assert all(np.isfinite(v) and v > 0 for v in t.values()), key
    assert table["assembly_ms"][(best["systems_per_thread"], best["column_chunk"])] == \
        min(table["assembly_ms"].values())
    assert table["op_gemv_ms"][(best["op_wg_target"], best["op_nt"])] == \
        min(table["op_gemv_ms"].values())
    assert table["ref_gemv_ms"][(best["ref_wg_target"], best["ref_nt"])] == \
        min(table["ref_gemv_ms"].values())
    sols = {"tuned": solve.sweep(pd, ctx=ctx)[0]}
    for label, cfg in (("8/256/256/0/4096/1", (8, 256, 256, 0, 4096, 1)),
                       ("32/128/4096/1/256/0", (32, 128, 4096, 1, 256, 0))):
        ctx.set_sweep_tuning(capi.C32, n, cfg)
        sols[label] = solve.sweep(pd, ctx=ctx)[0]
    ctx.set_sweep_tuning(capi.C32, n, None)
    for label, phi in sols.items():
        e = float(torch.linalg.vector_norm(phi - phi0) / torch.linalg.vector_norm(phi0))
        print(f"{label}: {e:.2e}")
        assert e < 1e-9, label
    ctx.close()


def test_sweep_assembly_invariant_to_tuning():
    """The batched fused assembly (ccab_assemble_system, the sweep kernel)
    with other systems per thread and column chunks: LHS entries equal to
    FP64 rounding (the phasor steps), right-hand sides (row sums over other
    chunks) equal to rounding."""
    from cc_abstemious_amd import influence_matrix
    pd = _sweep_case(level=3, nfreq=12, bc=capi.BC_NORMAL_VELOCITY, bcv=(0.7, -0.2))
    n = pd.get_num_eqn()
    _note(f"sweep assembly vs tuning n={n} nfreq=12")
    ctx = capi.Context()
    fs = pd.get_frequencies()
    rhs_inc = torch.zeros((len(fs), n), dtype=torch.complex128, device="cuda")
    l0, _, r0 = influence_matrix.assemble_system(pd, rhs_inc, fs, ctx=ctx)
    for cfg in ((8, 128, 1024, 1, 1024, 1), (32, 256, 1024, 1, 1024, 1)):
        ctx.set_sweep_tuning(capi.C64, n, cfg)
        l1, _, r1 = influence_matrix.assemble_system(pd, rhs_inc, fs, ctx=ctx)
        el = float((l1 - l0).abs().max() / l0.abs().max())
        er = float((r1 - r0).abs().max() / r0.abs().max())
        print(cfg, el, er)
        assert el < 1e-14 and er < 1e-13, cfg
    ctx.set_sweep_tuning(capi.C64, n, None)
    ctx.close()
