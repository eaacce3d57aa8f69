"""GPU parity: libccab (HIP, gfx950) against the CPU oracle, through the C ABI.

Bars: integer/index results exact; FP64 paths within the stated relative
tolerances of the FP64 oracle; complex-FP32 paths within FP32 tolerances; the
reference's own golden values where they apply.
"""
import math
import os

import numpy as np
import pytest
import torch

import oracle as O
from cc_abstemious_amd import capi, influence_matrix, incident_wave, mesh, solve
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver, colmajor

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
FP = np.array([[10.0, 0.0, 0.0]])
P_CLASSICAL = -3.9380091293979903e-05 + 4.7605686656319906e-04j
P_BM = 1.2125591274707629e-04 + 4.814465829556038e-04j


def _note(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def sphere():
    return mesh.read_vtk(os.path.join(GOLD, "sphere.vtk"), 3)


def _predata(m, method=capi.CLASSICAL, bc=capi.BC_NORMAL_VELOCITY, bcv=(0.0, 0.0), f=100.0,
             c=1500.0, rho=1000.0, solver=None, wave=None):
    return PreData(m, [f], c, rho, method, capi.EXTERIOR, bc, bcv,
                   wave or IncidentWave(capi.WAVE_PLANE, (1.0, 0.0, 0.0), (1.0, 0.0)),
                   solver or Solver())


def _oph(pd):
    return O.physics(pd.get_frequency(), pd.sound_speed, pd.mass_density, pd.method,
                     pd.problem_type, pd.bc_type, pd.bc_value)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_library_loaded_is_in_tree():
    L = capi.lib()
    assert os.path.dirname(os.path.abspath(L._name)) == os.path.join(ROOT, "cc_abstemious_amd")
    assert torch.cuda.is_available()


@pytest.mark.parametrize("method", [capi.CLASSICAL, capi.BURTON_MILLER])
def test_dense_surface_matrices_c128(sphere, method):
    _note(f"dense H,G c128 method={method} n=324")
    pd = _predata(sphere, method)
    H, G = influence_matrix.get_dense_surface_matrices(pd, capi.C128)
    Hr, Gr = O.dense_surface_matrices(_oph(pd), pd.cpts)
    assert _rel(H.cpu().numpy(), Hr) < 1e-13
    assert _rel(G.cpu().numpy(), Gr) < 1e-13


@pytest.mark.parametrize("method", [capi.CLASSICAL, capi.BURTON_MILLER])
def test_dense_surface_matrices_c64(sphere, method):
    _note(f"dense H,G c64 method={method}")
    pd = _predata(sphere, method, f=400.0)
    H, G = influence_matrix.get_dense_surface_matrices(pd, capi.C64)
    Hr, Gr = O.dense_surface_matrices(_oph(pd), pd.cpts)
    # entrywise FP32 accuracy relative to each matrix's scale
    assert _rel(H.cpu().numpy().astype(np.complex128), Hr) < 2e-6
    assert _rel(G.cpu().numpy().astype(np.complex128), Gr) < 2e-6


@pytest.mark.parametrize("bc,bcv", [(capi.BC_NORMAL_VELOCITY, (0.5, -0.2)),
                                    (capi.BC_PRESSURE, (0.3, 0.1)),
                                    (capi.BC_IMPEDANCE, (415.03, 83.006))])
def test_assemble_system_fused(sphere, bc, bcv):
    _note(f"assemble_system bc={bc}")
    pd = _predata(sphere, capi.BURTON_MILLER, bc, bcv, f=150.0, c=343.0, rho=1.21)
    rhs_inc = incident_wave.get_incident_surface(pd)
    l128, l64, rhs = influence_matrix.assemble_system(pd, rhs_inc, want_c128=True, want_c64=True)
    ph = _oph(pd)
    H, G = O.dense_surface_matrices(ph, pd.cpts)
    ri = O.incident_surface(ph, O.wave(O.WAVE_PLANE, (1, 0, 0), (1, 0)), pd.cpts)
    np.testing.assert_allclose(rhs_inc.cpu().numpy(), ri, rtol=1e-13, atol=1e-18)
    v = complex(*bcv)
    if bc == capi.BC_NORMAL_VELOCITY:
        L, R = H, G @ np.full(len(ri), v) - ri
    elif bc == capi.BC_PRESSURE:
        L, R = G, ri - H @ np.full(len(ri), v)
    else:
        f = (-1j * 2 * math.pi * 150.0 * 1.21) / v
        L, R = H + f * G, -ri
    assert _rel(l128[0].cpu().numpy(), L) < 1e-13
    assert _rel(l64[0].cpu().numpy().astype(complex), L) < 1e-6
    assert _rel(rhs[0].cpu().numpy(), R) < 1e-12


@pytest.mark.parametrize("spacing", ["uniform", "irregular"])
@pytest.mark.parametrize("bc,bcv", [(capi.BC_NORMAL_VELOCITY, (0.5, -0.2)),
                                    (capi.BC_PRESSURE, (0.3, 0.1))])
def test_assemble_system_sweep(sphere, spacing, bc, bcv):
    """One batched assembly over 11 frequencies (8 systems per thread sharing
    the pair geometry; with equally spaced wavenumbers e^{ikr} is stepped by
    e^{i dk r} from the first system of each group of 8) against one assembly
    per frequency and, for the last stepped system of the full group (7) and
    of the partial group (10), against the oracle."""
    if spacing == "uniform":
        ka = np.linspace(0.5, 6.0, 11)
    else:
        ka = np.array([0.5, 0.7, 1.3, 1.4, 2.9, 3.0, 3.7, 4.4, 5.2, 5.3, 6.0])
    fs = ka * 343.0 / (2 * math.pi)
    _note(f"batched assembly, {spacing} frequencies, bc={bc}")
    pd = _predata(sphere, capi.BURTON_MILLER, bc, bcv, f=fs[0], c=343.0, rho=1.21)
    n = pd.get_num_eqn()
    g = torch.Generator().manual_seed(3)
    rhs_inc = torch.complex(torch.randn(len(fs), n, generator=g, dtype=torch.float64),
                            torch.randn(len(fs), n, generator=g, dtype=torch.float64)).to(pd.device)
    l128, l64, rhs = influence_matrix.assemble_system(pd, rhs_inc, frequencies=fs,
                                                      want_c128=True, want_c64=True)
    worst = 0.0
    for b, f in enumerate(fs):
        s128, s64, srhs = influence_matrix.assemble_system(pd, rhs_inc[b:b + 1], frequencies=[f],
                                                           want_c128=True, want_c64=True)
        e = _rel(l128[b].cpu().numpy(), s128[0].cpu().numpy())
        worst = max(worst, e)
        assert e < 1e-14, (b, e)
        assert _rel(l64[b].cpu().numpy(), s64[0].cpu().numpy()) < 2e-7, b
        assert _rel(rhs[b].cpu().numpy(), srhs[0].cpu().numpy()) < 1e-13, b
    print(f"max entry difference batched vs per frequency: {worst:.2e}")
    v = complex(*bcv)
    for b in (7, 10):
        ph = O.physics(fs[b], 343.0, 1.21, capi.BURTON_MILLER, pd.problem_type, bc, bcv)
        H, G = O.dense_surface_matrices(ph, pd.cpts)
        L = H if bc == capi.BC_NORMAL_VELOCITY else G
        assert _rel(l128[b].cpu().numpy(), L) < 1e-13, b
        r = G @ np.full(n, v) if bc == capi.BC_NORMAL_VELOCITY else -(H @ np.full(n, v))
        ri = rhs_inc[b].cpu().numpy()
        expect = r - ri if bc == capi.BC_NORMAL_VELOCITY else r + ri
        assert _rel(rhs[b].cpu().numpy(), expect) < 1e-12, b


@pytest.mark.parametrize("side", [capi.LHS, capi.RHS])
def test_row_or_column(sphere, side):
    _note(f"row/column side={side}")
    pd = _predata(sphere, capi.BURTON_MILLER, capi.BC_PRESSURE, (1.0, 0.5))
    ph = _oph(pd)
    rows = [5]
    cols = [3, 5, 7, 5, 100, 0]          # diagonal twice: shift only at the first
    got = influence_matrix.get_surface_row_or_column(pd, rows, cols, side).cpu().numpy()
    ref = O.surface_row_or_column(ph, pd.cpts, side, rows, cols)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-16)
    got = influence_matrix.get_surface_row_or_column(pd, cols, [5], side).cpu().numpy()
    ref = O.surface_row_or_column(ph, pd.cpts, side, cols, [5])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-16)
    with pytest.raises(capi.CcabError):
        influence_matrix.get_surface_row_or_column(pd, [1, 2], [3, 4], side)


def test_solve_lu_reference_fixture():
    # solve.rs:211-218 through the GPU LU (c128, FP64 MFMA)
    _note("solve_lu 100x100 fixture")
    a, b = O.generate_random_ab(100, 100, 10)
    A = colmajor(100, 100, torch.complex128, "cuda").copy_(torch.from_numpy(a))
    x = torch.from_numpy(b.copy()).cuda()
    solve.solve_lu(A, x)
    x0 = x[0].item()
    assert abs(x0.real - 0.0035386336544537098) < 1e-14
    assert abs(x0.imag - 0.0017794290945383035) < 1e-14
    np.testing.assert_allclose(x.cpu().numpy(), O.lu_solve(a, b), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("n,batch,prec", [(64, 1, capi.C128), (200, 3, capi.C128),
                                          (333, 2, capi.C64), (700, 2, capi.C128)])
def test_lu_batched_vs_oracle(n, batch, prec):
    _note(f"lu batched n={n} batch={batch} prec={prec}")
    rng = np.random.default_rng(n)
    dt = torch.complex128 if prec == capi.C128 else torch.complex64
    As = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(batch)]
    bs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(batch)]
    A = torch.stack([torch.from_numpy(a.T.copy()) for a in As]).to(dt).cuda().transpose(1, 2)
    B = torch.stack([torch.from_numpy(b) for b in bs]).to(dt).cuda()
    ipiv = solve.lu_factor_batched(A)
    solve.lu_solve_batched(A, ipiv, B)
    tol = 1e-10 if prec == capi.C128 else 5e-3
    for k in range(batch):
        ref = O.lu_solve(As[k], bs[k])
        got = B[k].cpu().numpy().astype(np.complex128)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol
        if prec == capi.C128:
            # same pivot sequence as the oracle (nalgebra rule) -> identical pivots
            W = np.array(As[k], dtype=np.complex128, order="F")
            piv = np.zeros(n, np.int64)
            O.lib().orc_lu_factor(n, O._p(W), n, O._p(piv))
            assert (ipiv[k].cpu().numpy() == piv).all()


@pytest.mark.parametrize("prec", [capi.C64, capi.C128])
@pytest.mark.parametrize("M,N,K", [(64, 64, 64), (130, 77, 64), (257, 300, 33), (5, 7, 3)])
def test_cgemm_update(prec, M, N, K):
    _note(f"cgemm prec={prec} {M}x{N}x{K}")
    rng = np.random.default_rng(M * 7 + N)
    dt = torch.complex128 if prec == capi.C128 else torch.complex64
    a = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    b = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    c = rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))
    A = colmajor(M, K, dt, "cuda").copy_(torch.from_numpy(a))
    B = colmajor(K, N, dt, "cuda").copy_(torch.from_numpy(b))
    C = colmajor(M, N, dt, "cuda").copy_(torch.from_numpy(c))
    ctx = capi.default_context()
    ctx.call("ccab_cgemm_update", prec, M, N, K, solve._ptr(A), M, solve._ptr(B), K, solve._ptr(C), M)
    ctx.synchronize()
    ref = c - a @ b
    got = C.cpu().numpy().astype(np.complex128)
    tol = 1e-12 if prec == capi.C128 else 1e-5
    assert _rel(got, ref) < tol


def test_mixed_precision_refinement():
    _note("mixed c64 LU + FP64 IR n=500 batch=3")
    rng = np.random.default_rng(1)
    n, nb = 500, 3
    As = [np.eye(n) * 6 + rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
          for _ in range(nb)]
    bs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(nb)]
    S = torch.stack([torch.from_numpy(a.T.copy()) for a in As]).cuda()
    A128 = S.transpose(1, 2)
    A64 = S.to(torch.complex64).transpose(1, 2)
    x = torch.stack([torch.from_numpy(b) for b in bs]).cuda()
    x, st = solve.solve_mixed_batched(A128, A64, x, 1e-13, 20)
    for k in range(nb):
        ref = O.lu_solve(As[k], bs[k])
        err = np.linalg.norm(x[k].cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err < 1e-10, err
        assert st[k]["ir_steps"] >= 1 and st[k]["rel_residual"] < 1e-12


@pytest.mark.parametrize("ref_mgs", [True, False])
def test_gmres_rng_fixture_and_counts(ref_mgs):
    _note(f"gmres fixture ref_mgs={ref_mgs}")
    a, b = O.generate_random_ab(100, 100, 10)
    A = colmajor(100, 100, torch.complex128, "cuda").copy_(torch.from_numpy(a))
    x = torch.from_numpy(b.copy()).cuda()
    gm = solve.GMRES(100, 1e-8, ref_mgs=ref_mgs)
    gm.a = A
    gm.solve(x)
    x0 = x[0].item()
    assert abs(x0.real - 0.0035386336544537098) < 1e-6
    assert abs(x0.imag - 0.0017794290945383035) < 1e-6
    xr, info = O.gmres(a, b, 100, 1e-8, ref_mgs=ref_mgs)
    # same iteration and matvec counts as the oracle's restatement of gmres.rs
    assert gm.stats["num_mv"] == info["num_mv"]
    assert gm.stats["iterations"] == info["iterations"]
    np.testing.assert_allclose(x.cpu().numpy(), xr, rtol=1e-9, atol=1e-12)


def _field_from(pd, phi, vn):
    ph = _oph(pd)
    w = O.wave(O.WAVE_PLANE, (1, 0, 0), (1, 0))
    return O.field(ph, pd.cpts, FP, phi, vn, O.incident_field(ph, w, FP))[0]


@pytest.mark.parametrize("method,kind,expect,tol", [
    (capi.CLASSICAL, capi.SOLVER_DIRECT, P_CLASSICAL, 1e-12),
    (capi.CLASSICAL, capi.SOLVER_ITERATIVE, P_CLASSICAL, None),
    (capi.BURTON_MILLER, capi.SOLVER_DIRECT, P_BM, 1e-12),
    (capi.CLASSICAL, capi.SOLVER_DIRECT_MIXED, P_CLASSICAL, 1e-10),
    (capi.BURTON_MILLER, capi.SOLVER_DIRECT_MIXED, P_BM, 1e-10),
    (capi.CLASSICAL, capi.SOLVER_ITERATIVE_MIXED, P_CLASSICAL, 1e-9),
    (capi.BURTON_MILLER, capi.SOLVER_ITERATIVE_MIXED, P_BM, 1e-9),
])
def test_reference_fixtures_end_to_end(sphere, method, kind, expect, tol):
    """verification.rs:95-205 through the GPU surface solve."""
    _note(f"fixture method={method} solver={kind}")
    gtol = 1e-5 if kind == capi.SOLVER_ITERATIVE else 1e-12
    pd = _predata(sphere, method, solver=Solver(kind=kind, tolerance=gtol, max_iterations=1000))
    rhs_inc = incident_wave.get_incident_surface(pd)
    phi, vn = solve.solve_for_surface(pd, rhs_inc)
    p = _field_from(pd, phi.cpu().numpy(), vn.cpu().numpy())
    if tol is None:   # GMRES fixture: abs 1e-8 (verification.rs:146-147)
        assert abs(p.real - expect.real) < 1e-8 and abs(p.imag - expect.imag) < 1e-8
    else:
        assert abs(p - expect) / abs(expect) < tol


def test_sweep_mixed_matches_fp64():
    _note("sweep mixed vs c128, icosphere level 2, 6 frequencies, BM")
    m = mesh.icosphere(2)
    ka = np.linspace(0.5, 3.0, 6)
    fs = ka * 343.0 / (2 * math.pi)
    pd = PreData(m, fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY,
                 (0, 0), IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_DIRECT_MIXED))
    phi_m, vn_m, st = solve.sweep(pd)
    pd.solver = Solver(kind=capi.SOLVER_DIRECT)
    phi_d, vn_d, _ = solve.sweep(pd)
    w = O.wave(O.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0))
    for k, f in enumerate(fs):
        ph = O.physics(f, 343.0, 1.21, O.BURTON_MILLER)
        rhs = O.incident_surface(ph, w, pd.cpts)
        phi_o, _, _ = O.solve_surface_dense(ph, pd.cpts, rhs, O.SOLVER_DIRECT)
        e_m = np.linalg.norm(phi_m[k].cpu().numpy() - phi_o) / np.linalg.norm(phi_o)
        e_d = np.linalg.norm(phi_d[k].cpu().numpy() - phi_o) / np.linalg.norm(phi_o)
        assert e_d < 1e-12 and e_m < 1e-10, (k, e_d, e_m)


def test_residual_matrix_needs_fp16_operator():
    _note("residual_matrix C32 with a c64 operator is refused")
    pd = _predata(mesh.icosphere(1), capi.BURTON_MILLER,
                  solver=Solver(kind=capi.SOLVER_ITERATIVE_MIXED, matrix_prec=capi.C64,
                                residual_matrix=capi.C32))
    with pytest.raises(capi.CcabError, match="residual_matrix"):
        solve.sweep(pd)
    pd.solver = Solver(kind=capi.SOLVER_DIRECT, residual_matrix=capi.C32)
    with pytest.raises(capi.CcabError, match="residual_matrix"):
        solve.sweep(pd)


def test_error_paths():
    _note("error paths")
    ctx = capi.default_context()
    ph = capi.physics(-1.0)
    d = torch.zeros((4, 8), dtype=torch.float64, device="cuda")
    out = torch.zeros(4, dtype=torch.complex128, device="cuda")
    w = capi.wave()
    with pytest.raises(capi.CcabError) as e:
        ctx.call("ccab_incident_surface", __import__("ctypes").byref(ph),
                 __import__("ctypes").byref(w), solve._ptr(d), 4, solve._ptr(out))
    assert e.value.code == capi.EINVAL
    # singular matrix -> ESINGULAR
    A = torch.zeros((1, 8, 8), dtype=torch.complex128, device="cuda").transpose(1, 2)
    with pytest.raises(capi.CcabError) as e:
        solve.lu_factor_batched(A)
    assert e.value.code == capi.ESINGULAR


@pytest.mark.parametrize("n,prec", [(2900, capi.C128), (5200, capi.C128), (3000, capi.C64)])
def test_lu_large_panels_vs_lapack(n, prec):
    """Sizes that exercise every register-slab panel width (rows per lane 6 and
    11) and the XCD tile order; checked against LAPACK zgesv (numpy)."""
    _note(f"lu large n={n} prec={prec}")
    rng = np.random.default_rng(7)
    dt = torch.complex128 if prec == capi.C128 else torch.complex64
    shift = 3.0 if prec == capi.C128 else 2.0 * math.sqrt(2 * n)   # keep c64 well conditioned
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + shift * np.eye(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    A = torch.from_numpy(a.T.copy()).to(dt).cuda().unsqueeze(0).transpose(1, 2)
    B = torch.from_numpy(b).to(dt).cuda().unsqueeze(0).contiguous()
    ipiv = solve.lu_factor_batched(A)
    solve.lu_solve_batched(A, ipiv, B)
    ref = np.linalg.solve(a, b)
    err = np.linalg.norm(B[0].cpu().numpy().astype(complex) - ref) / np.linalg.norm(ref)
    assert err < (1e-10 if prec == capi.C128 else 1e-5), err


@pytest.mark.parametrize("x0_kind", ["zero", "warm"])
def test_sweep_gmres_ir_matches_fp64(x0_kind):
    _note(f"sweep GMRES-IR vs c128 LU, icosphere level 2, BM, x0={x0_kind}")
    m = mesh.icosphere(2)
    ka = np.linspace(0.5, 4.0, 7)
    fs = ka * 343.0 / (2 * math.pi)
    pd = PreData(m, fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY,
                 (0, 0), IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_DIRECT))
    phi_d, vn_d, _ = solve.sweep(pd)
    pd.solver = Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-12, restart=30,
                       ir_tolerance=1e-4, ir_max_steps=30, ref_mgs=False)
    x0 = None
    if x0_kind == "warm":   # a perturbed solution as the initial guess
        x0 = phi_d * (1 + 0.01 * torch.randn_like(phi_d.real))
    phi_g, vn_g, st = solve.sweep(pd, x0=x0)
    for k in range(len(fs)):
        e = float(torch.linalg.vector_norm(phi_g[k] - phi_d[k]) / torch.linalg.vector_norm(phi_d[k]))
        assert e < 1e-9, (k, e, st[k])
        assert st[k]["converged"] == 1 and st[k]["rel_residual"] <= 1e-12
    if x0_kind == "warm":
        pd.solver.kind = capi.SOLVER_ITERATIVE_MIXED
        _, _, st0 = solve.sweep(pd)
        assert sum(s["iterations"] for s in st) < sum(s["iterations"] for s in st0)



@pytest.mark.parametrize("level,residual", [(2, capi.C128), (3, capi.C128), (3, capi.C32)])
def test_sweep_gmres_ir_fp16_operator_matches_fp64(level, residual):
    """The complex-FP16 (scaled) GMRES operator inside FP64 refinement reaches
    the FP64 LU solution; only the inner solve sees FP16 entries.  With
    residual C32 the refinement's matrix is the FP16 operator plus its
    complex-FP32 remainder (about 2^-35 relative per entry; without the
    remainder the error would stay near 1e-4), and its residual is carried
    as r - Q (Hbar y) - A_l d between cycles: the true residual against the
    c128 matrix must agree with the reported one."""
    _note(f"sweep GMRES-IR with c32 operator vs c128 LU, icosphere level {level}, BM, "
          f"residual matrix {residual}")
    m = mesh.icosphere(level)
    ka = np.linspace(0.5, 8.0, 6)
    fs = ka * 343.0 / (2 * math.pi)
    pd = PreData(m, fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY,
                 (0, 0), IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_DIRECT))
    phi_d, vn_d, _ = solve.sweep(pd)
    pd.solver = Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-10, restart=50,
                       ir_tolerance=1e-3, ir_max_steps=30, ref_mgs=False, matrix_prec=capi.C32,
                       residual_matrix=residual)
    phi_h, _, st = solve.sweep(pd)
    for k in range(len(fs)):
        e = float(torch.linalg.vector_norm(phi_h[k] - phi_d[k]) / torch.linalg.vector_norm(phi_d[k]))
        assert e < 1e-8, (k, e, st[k])
        assert st[k]["converged"] == 1 and st[k]["rel_residual"] <= 1e-10
    print("iterations", [s["iterations"] for s in st], "cycles", [s["ir_steps"] for s in st])
    if residual == capi.C32:
        for k in (0, len(fs) - 1):
            pk = PreData(m, [fs[k]], 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR,
                         capi.BC_NORMAL_VELOCITY, (0, 0),
                         IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)), Solver())
            rhs_inc = incident_wave.get_incident_surface(pk)
            l128, _, rhs = influence_matrix.assemble_system(pk, rhs_inc)
            r = l128[0] @ phi_h[k] - rhs[0]
            true_rel = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(rhs[0]))
            print(f"frequency {k}: reported {st[k]['rel_residual']:.3e} true {true_rel:.3e}")
            # tolerance 1e-10 plus the remainder's 2^-35 representation error
            assert true_rel <= 3e-10
        # warm start: the first residual b - A x0 with A x0 formed by the
        # assembly from the FP64 entries, then carried as above
        g = torch.Generator().manual_seed(11)
        x0 = phi_d * (1 + 0.01 * torch.randn(phi_d.shape, generator=g, dtype=torch.float64).to(phi_d.device))
        phi_w, _, st_w = solve.sweep(pd, x0=x0)
        for k in range(len(fs)):
            e = float(torch.linalg.vector_norm(phi_w[k] - phi_d[k]) / torch.linalg.vector_norm(phi_d[k]))
            assert e < 1e-8, (k, e, st_w[k])
            assert st_w[k]["converged"] == 1 and st_w[k]["rel_residual"] <= 1e-10
        assert sum(s["iterations"] for s in st_w) < sum(s["iterations"] for s in st)
        pk = PreData(m, [fs[-1]], 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR,
                     capi.BC_NORMAL_VELOCITY, (0, 0),
                     IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)), Solver())
        l128, _, rhs = influence_matrix.assemble_system(pk, incident_wave.get_incident_surface(pk))
        r = l128[0] @ phi_w[-1] - rhs[0]
        true_rel = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(rhs[0]))
        print(f"warm start: iterations {sum(s['iterations'] for s in st_w)} vs "
              f"{sum(s['iterations'] for s in st)}; true residual {true_rel:.3e}")
        assert true_rel <= 3e-10


def test_sweep_fp16_operator_n5120():
    """n = 5120, where the Gram-Schmidt step runs as the multi-workgroup
    kernel (grid barriers between its partial dot products): the
    FP16-operator GMRES-IR sweep reaches the FP64 LU solution to the
    refinement's tolerance, with the same iteration counts on a repeat
    (the kernel's fixed-order sums make runs reproducible)."""
    _note("sweep n=5120, FP16-operator GMRES-IR (multi-workgroup Gram-Schmidt) vs FP64 LU")
    m = mesh.icosphere(4)
    fs = np.array([1.0, 3.5, 6.0]) * 343.0 / (2 * math.pi)
    pd = PreData(m, fs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR, capi.BC_NORMAL_VELOCITY,
                 (0, 0), IncidentWave(capi.WAVE_PLANE, (0.2, 0.3, 0.9), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_DIRECT))
    phi_d, _, _ = solve.sweep(pd)
    pd.solver = Solver(kind=capi.SOLVER_ITERATIVE_MIXED, tolerance=1e-7, restart=30,
                       ir_tolerance=1e-3, ir_max_steps=20, ref_mgs=False, matrix_prec=capi.C32,
                       residual_matrix=capi.C32)
    runs = []
    for _ in range(2):
        phi, _, st = solve.sweep(pd)
        runs.append((phi.cpu().numpy(), [s_["iterations"] for s_ in st]))
        for k in range(len(fs)):
            e = float(np.linalg.norm(runs[-1][0][k] - phi_d[k].cpu().numpy())
                      / np.linalg.norm(phi_d[k].cpu().numpy()))
            assert st[k]["converged"] == 1 and e < 1e-6, (k, e)
    print("iterations", runs[0][1])
    assert runs[0][1] == runs[1][1]
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
