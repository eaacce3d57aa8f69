"""GPU parity of the hierarchical-matrix path (src/solve/h_matrix.rs, aca.rs,
h_matrix/cluster.rs, h_matrix/block.rs, solve.rs:74-150) against the oracle
and the reference's H-matrix fixture (verification.rs:174-175).

Both sides take the same ACA pivots (ChaCha12 reference-row draws, seed 10),
so the GPU blocks must agree with the oracle's to rounding, not only to the
ACA tolerance."""
import os

import numpy as np
import pytest
import torch

import oracle as O
from cc_abstemious_amd import capi, h_matrix, incident_wave, mesh, solve
from cc_abstemious_amd.preprocess import IncidentWave, PreData, Solver

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
FP = np.array([[10.0, 0.0, 0.0]])
P_HMATRIX = -3.9380308176032198e-05 + 4.7605563272935873e-04j   # verification.rs:174-175


def _note(msg):
    print(f"[case] {msg}", flush=True)


def _predata(m, method=capi.CLASSICAL, bc=capi.BC_NORMAL_VELOCITY, bcv=(0.0, 0.0), f=100.0,
             c=1500.0, rho=1000.0, solver=None):
    return PreData(m, [f], c, rho, method, capi.EXTERIOR, bc, bcv,
                   IncidentWave(capi.WAVE_PLANE, (1.0, 0.0, 0.0), (1.0, 0.0)),
                   solver or Solver(kind=capi.SOLVER_HIERARCHICAL, tolerance=1e-5,
                                    max_iterations=1000))


def _oph(pd):
    return O.physics(pd.get_frequency(), pd.sound_speed, pd.mass_density, pd.method,
                     pd.problem_type, pd.bc_type, pd.bc_value)


def _oracle_h(pd, side, leaf=32, tol=1e-4, eta=4.0):
    ph = _oph(pd)
    cp = pd.cpts
    return O.HMatrix(cp.shape[0], lambda r, c: O.surface_row_or_column(ph, cp, side, r, c).reshape(-1),
                     cp, leaf, tol, eta)


def _lumpy(nu, seed=7):
    """Geodesic sphere (20 nu^2 triangles) with a random low-order radial
    perturbation: no mesh symmetry, so no two ACA pivot candidates tie and
    the GPU and the oracle must take the same pivots."""
    c = np.random.default_rng(seed).standard_normal(81) * 0.03
    c[0] = 0.0
    return mesh.perturbed_sphere(nu, c)


@pytest.fixture(scope="module")
def sphere():
    return mesh.read_vtk(os.path.join(GOLD, "sphere.vtk"), 3)


def test_hmatrix_fixture_end_to_end(sphere):
    """verification.rs:174-175: hierarchical solve of the sphere fixture."""
    _note("H-matrix fixture, sphere.vtk n=324, classical, GMRES tol 1e-5")
    pd = _predata(sphere)
    rhs_inc = incident_wave.get_incident_surface(pd)
    phi, vn = solve.solve_for_surface(pd, rhs_inc)
    ph = _oph(pd)
    w = O.wave(O.WAVE_PLANE, (1, 0, 0), (1, 0))
    p = O.field(ph, pd.cpts, FP, phi.cpu().numpy(), vn.cpu().numpy(),
                O.incident_field(ph, w, FP))[0]
    assert abs(p.real - P_HMATRIX.real) < 1e-10 and abs(p.imag - P_HMATRIX.imag) < 1e-10
    # the oracle takes the same GMRES path: same matvec count, same solution
    phi_o, vn_o, st_o = O.solve_surface_hmatrix(ph, pd.cpts, rhs_inc.cpu().numpy(), 1000, 1e-5,
                                                32, 1e-4, 4.0)
    assert pd.last_stats["num_mv"] == st_o[0]
    assert pd.last_stats["iterations"] == st_o[1]
    np.testing.assert_allclose(phi.cpu().numpy(), phi_o, rtol=0, atol=1e-9 * np.abs(phi_o).max())


@pytest.mark.parametrize("nu,method,side", [(4, capi.CLASSICAL, capi.LHS),
                                            (8, capi.BURTON_MILLER, capi.LHS),
                                            (8, capi.BURTON_MILLER, capi.RHS),
                                            (8, capi.CLASSICAL, capi.RHS)])
def test_hmatrix_blocks_vs_oracle(nu, method, side):
    """Block structure, ranks, norm and the represented matrix."""
    m = _lumpy(nu)
    pd = _predata(m, method, capi.BC_PRESSURE, (1.0, 0.5), f=400.0, c=343.0, rho=1.21)
    n = pd.get_num_eqn()
    _note(f"H blocks vs oracle icosphere n={n} method={method} side={side}")
    H = h_matrix.HMatrix.new_from(pd, side)
    Ho = _oracle_h(pd, side)
    st, sto = H.stats(), Ho.stats()
    assert st["dense_fallback"] == 0
    for k in ("admissible", "inadmissible", "storage", "rank_sum"):
        assert st[k] == sto[k], k
    # the norm estimate (aca.rs:212-241) uses unconjugated dots, so it is not
    # invariant under rotations of nearly equal singular vectors: 1e-8 relative
    assert abs(st["norm"] - Ho.norm()) <= 1e-8 * Ho.norm(), (st["norm"], Ho.norm())
    full = H.to_full().cpu().numpy()
    full_o = Ho.to_full()
    e = np.linalg.norm(full - full_o) / np.linalg.norm(full_o)
    print(f"to_full vs oracle: {e:.3e}")
    assert e < 1e-10
    # and it approximates the dense side matrix to the ACA tolerance
    ph, allc = _oph(pd), list(range(n))
    A = np.stack([O.surface_row_or_column(ph, pd.cpts, side, [i], allc) for i in range(n)])
    assert np.linalg.norm(full - A) / np.linalg.norm(A) < 1e-3


def test_hmatrix_symmetric_mesh_within_aca_tolerance():
    """On an icosphere many pivot candidates tie exactly; the last bit of the
    kernel values decides them, so the GPU and the oracle may take different
    (equally valid) pivots.  Both must then approximate the same matrix to
    the ACA tolerance."""
    m = mesh.icosphere(3)
    pd = _predata(m, capi.BURTON_MILLER, capi.BC_PRESSURE, (1.0, 0.5), f=400.0, c=343.0,
                  rho=1.21)
    n = pd.get_num_eqn()
    _note(f"H icosphere n={n} within ACA tolerance")
    H = h_matrix.HMatrix.new_from(pd, capi.RHS)
    Ho = _oracle_h(pd, capi.RHS)
    full, full_o = H.to_full().cpu().numpy(), Ho.to_full()
    ph, allc = _oph(pd), list(range(n))
    A = np.stack([O.surface_row_or_column(ph, pd.cpts, capi.RHS, [i], allc) for i in range(n)])
    e_gpu = np.linalg.norm(full - A) / np.linalg.norm(A)
    e_orc = np.linalg.norm(full_o - A) / np.linalg.norm(A)
    e_x = np.linalg.norm(full - full_o) / np.linalg.norm(A)
    print(f"vs dense: gpu {e_gpu:.3e} oracle {e_orc:.3e}; gpu vs oracle {e_x:.3e}")
    assert e_gpu < 1e-3 and e_orc < 1e-3 and e_x < 1e-4
    assert e_gpu < 3.0 * e_orc + 1e-12


def test_hmatrix_gemv_vs_oracle():
    m = _lumpy(8)
    pd = _predata(m, capi.BURTON_MILLER, f=300.0, c=343.0, rho=1.21)
    n = pd.get_num_eqn()
    _note(f"H gemv vs oracle lumpy sphere n={n}")
    H = h_matrix.HMatrix.new_from(pd, capi.LHS)
    Ho = _oracle_h(pd, capi.LHS)
    assert H.stats()["dense_fallback"] == 0
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    alpha, beta = 0.3 - 1.1j, 0.5 + 0.25j
    dev = pd.device
    xt = torch.from_numpy(x).to(dev)
    for bet in (beta, 1.0, 0.0):
        yt = torch.from_numpy(y0.copy()).to(dev)
        H.gemv(alpha, xt, bet, yt)
        ref = Ho.gemv(alpha, x, bet, y0)
        e = np.linalg.norm(yt.cpu().numpy() - ref) / np.linalg.norm(ref)
        print(f"gemv beta={bet}: {e:.3e}")
        assert e < 1e-10
    # gemv agrees with the matrix it represents
    full = H.to_full()
    yt = torch.zeros(n, dtype=torch.complex128, device=dev)
    H.gemv(1.0, xt, 0.0, yt)
    ref = (full @ xt).cpu().numpy()
    e = np.linalg.norm(yt.cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"gemv vs to_full: {e:.3e}")
    assert e < 1e-13


@pytest.mark.parametrize("bc,bcv", [(capi.BC_PRESSURE, (1.0, 0.5)),
                                    (capi.BC_NORMAL_VELOCITY, (0.5, -0.2)),
                                    (capi.BC_IMPEDANCE, (400.0, 30.0))])
def test_hierarchical_solve_all_bcs(bc, bcv):
    """get_surface_hmatrix (solve.rs:74-150) incl. the RHS-side H matrix."""
    m = mesh.icosphere(2)
    pd = _predata(m, capi.BURTON_MILLER, bc, bcv, f=300.0, c=343.0, rho=1.21)
    n = pd.get_num_eqn()
    _note(f"hierarchical solve bc={bc} n={n}")
    rhs_inc = incident_wave.get_incident_surface(pd)
    phi, vn = solve.solve_for_surface(pd, rhs_inc)
    phi_o, vn_o, st_o = O.solve_surface_hmatrix(_oph(pd), pd.cpts, rhs_inc.cpu().numpy(), 1000,
                                                1e-5, 32, 1e-4, 4.0)
    assert pd.last_stats["num_mv"] == st_o[0]
    for got, ref in ((phi, phi_o), (vn, vn_o)):
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=0, atol=1e-9 * np.abs(ref).max())


def test_hierarchical_sweep_matches_dense():
    """Sweep over frequencies with the H matrix; compare with the FP64 LU."""
    m = mesh.icosphere(3)
    ka = np.array([1.0, 2.5, 4.0])
    freqs = ka * 343.0 / (2 * np.pi)
    pd = PreData(m, freqs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR,
                 capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                 IncidentWave(capi.WAVE_PLANE, (0.0, 0.6, 0.8), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_HIERARCHICAL, tolerance=1e-8, max_iterations=1000,
                        aca_tolerance=1e-6))
    _note(f"hierarchical sweep n={pd.get_num_eqn()} nfreq=3")
    phi_h, _, st = solve.sweep(pd)
    pd.solver = Solver(kind=capi.SOLVER_DIRECT)
    phi_d, _, _ = solve.sweep(pd)
    err = (torch.linalg.vector_norm(phi_h - phi_d, dim=1) / torch.linalg.vector_norm(phi_d, dim=1))
    assert float(err.max()) < 1e-4
    assert all(s["converged"] for s in st)


def test_hmatrix_fp32_storage():
    """Blocks built in FP64 and stored in complex FP32 (ccab_hmatrix_build_ex,
    storage C64): the represented matrix equals the FP64-stored one to FP32
    rounding (entries within 2^-24 of the largest), the matvec agrees with it,
    the stored storage statistics are unchanged, and the hierarchical sweep
    with matrix_prec C64 reaches the FP64-storage solution."""
    m = _lumpy(8)
    pd = _predata(m, capi.BURTON_MILLER, f=300.0, c=343.0, rho=1.21)
    n = pd.get_num_eqn()
    _note(f"H matrix FP32 storage lumpy sphere n={n}")
    H64 = h_matrix.HMatrix.new_from(pd, capi.LHS)
    H32 = h_matrix.HMatrix.new_from(pd, capi.LHS, storage=capi.C64)
    s64, s32 = H64.stats(), H32.stats()
    assert s64 == s32
    f64, f32 = H64.to_full(), H32.to_full()
    e = float((f32 - f64).abs().max() / f64.abs().max())
    print(f"to_full FP32 vs FP64 storage: {e:.3e}")
    assert 0.0 < e < 2.0 ** -23
    rng = np.random.default_rng(9)
    xt = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(pd.device)
    y = torch.zeros(n, dtype=torch.complex128, device=pd.device)
    H32.gemv(0.7 - 0.2j, xt, 0.0, y)
    ref = (0.7 - 0.2j) * (f32 @ xt)
    e = float(torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref))
    print(f"FP32-storage gemv vs its to_full: {e:.3e}")
    assert e < 1e-13
    # the sweep
    freqs = np.array([1.5, 3.0]) * 343.0 / (2 * np.pi)
    sp = PreData(mesh.icosphere(3), freqs, 343.0, 1.21, capi.BURTON_MILLER, capi.EXTERIOR,
                 capi.BC_NORMAL_VELOCITY, (0.0, 0.0),
                 IncidentWave(capi.WAVE_PLANE, (0.0, 0.6, 0.8), (1.0, 0.0)),
                 Solver(kind=capi.SOLVER_HIERARCHICAL, tolerance=1e-8, max_iterations=1000,
                        aca_tolerance=1e-6))
    p64, _, st64 = solve.sweep(sp)
    sp.solver.matrix_prec = capi.C64
    p32, _, st32 = solve.sweep(sp)
    err = torch.linalg.vector_norm(p32 - p64, dim=1) / torch.linalg.vector_norm(p64, dim=1)
    print("sweep FP32 vs FP64 storage", err.tolist(), [s["num_mv"] for s in st64],
          [s["num_mv"] for s in st32])
    assert float(err.max()) < 1e-6
    assert all(s["converged"] for s in st32)


def test_gmres_hmatrix_class_api():
    """GMRES(max_it, thresh) with .hmatrix set (gmres.rs:18, solve.rs:119-125)."""
    m = mesh.icosphere(2)
    pd = _predata(m, capi.BURTON_MILLER, f=200.0, c=343.0, rho=1.21)
    n = pd.get_num_eqn()
    _note(f"GMRES.hmatrix n={n}")
    H = h_matrix.HMatrix.new_from(pd, capi.LHS)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    gm = solve.GMRES(1000, 1e-10)
    gm.hmatrix = H
    x = torch.from_numpy(b.copy()).to(pd.device)
    gm.solve(x)
    full = H.to_full()
    r = torch.from_numpy(b).to(pd.device) - full @ x
    # backward error with the H norm estimate (gmres.rs:142-144)
    be = float(torch.linalg.vector_norm(r)) / (H.get_norm() * float(torch.linalg.vector_norm(x))
                                               + np.linalg.norm(b))
    print(f"backward error {be:.3e}, lib {gm.stats['backward_error']:.3e}")
    assert be < 1.01e-10
    assert gm.stats["converged"] == 1


def test_hmatrix_errors():
    m = mesh.icosphere(1)
    pd = _predata(m)
    ctx = capi.default_context()
    import ctypes as ct
    h = ct.c_void_p()
    with pytest.raises(capi.CcabError):
        ctx.call("ccab_hmatrix_build", ct.byref(pd.physics()), ct.c_void_p(pd.d_cpts.data_ptr()),
                 pd.get_num_eqn(), 7, 32, 4.0, 1e-4, ct.byref(h))
    with pytest.raises(capi.CcabError):
        ctx.call("ccab_hmatrix_build", ct.byref(pd.physics()), ct.c_void_p(pd.d_cpts.data_ptr()),
                 pd.get_num_eqn(), capi.LHS, 0, 4.0, 1e-4, ct.byref(h))


@pytest.mark.parametrize("pl", ["0", "1", "4"])
def test_aca_wavefront_variants_agree(pl, monkeypatch):
    """The ACA of blocks up to 64, 128 (the default) or 256 rows and columns
    on one wavefront each (lanes holding the residual entries, no workgroup
    barrier) against a workgroup per block (CCAB_HM_WAVE_ACA=0) and against
    the default: the same blocks, ranks and represented matrix (pivots are
    the same; only the norm estimate's sums are formed in another order).
    Leaf 64, eta 1 on 5120 lumpy triangles: blocks from 32 to over 256 rows."""
    m = _lumpy(16)
    pd = _predata(m, capi.BURTON_MILLER, f=2000.0, c=343.0, rho=1.21)
    _note(f"ACA wavefront variant CCAB_HM_WAVE_ACA={pl} vs default, n={pd.get_num_eqn()}")
    ref = h_matrix.HMatrix.new_from(pd, capi.LHS, 64, 1e-4, 1.0)
    monkeypatch.setenv("CCAB_HM_WAVE_ACA", pl)
    H = h_matrix.HMatrix.new_from(pd, capi.LHS, 64, 1e-4, 1.0)
    st, st0 = H.stats(), ref.stats()
    for k in ("admissible", "inadmissible", "storage", "rank_sum", "dense_fallback"):
        assert st[k] == st0[k], k
    assert abs(st["norm"] - st0["norm"]) <= 1e-12 * st0["norm"]
    full, full0 = H.to_full(), ref.to_full()
    e = float(torch.linalg.vector_norm(full - full0) / torch.linalg.vector_norm(full0))
    print(f"to_full vs default: {e:.3e}")
    assert e < 1e-13
