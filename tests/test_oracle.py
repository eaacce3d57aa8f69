"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each test is anchored on something the oracle formula cannot simply restate:
the reference's own golden values (src/tests/verification.rs, solve.rs,
aca.rs, h_matrix.rs, cluster.rs, block.rs, mesh_data.rs tests), closed forms,
LAPACK, truncated SVDs and invariants.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SPHERE = os.path.join(GOLD, "sphere.vtk")

# verification.rs:95-205 common setup
FP = np.array([[10.0, 0.0, 0.0]])
P_CLASSICAL = -3.9380091293979903e-05 + 4.7605686656319906e-04j   # verification.rs:117-119
P_HMATRIX = -3.9380308176032198e-05 + 4.7605563272935873e-04j     # verification.rs:174-175
P_BM = 1.2125591274707629e-04 + 4.814465829556038e-04j            # verification.rs:202-204


@pytest.fixture(scope="module")
def sphere_cpts():
    return O.mesh_cpts_from_vtk(SPHERE, 3)


def _fixture_field(cp, method, solver, hier=False):
    ph = O.physics(100.0, 1500.0, 1000.0, method, O.EXTERIOR, O.BC_NORMAL_VELOCITY, (0, 0))
    w = O.wave(O.WAVE_PLANE, (1, 0, 0), (1, 0))
    rhs = O.incident_surface(ph, w, cp)
    if hier:
        phi, vn, _ = O.solve_surface_hmatrix(ph, cp, rhs, 1000, 1e-5, 32, 1e-4, 4.0)
    else:
        phi, vn, _ = O.solve_surface_dense(ph, cp, rhs, solver, 1000, 1e-5)
    return O.field(ph, cp, FP, phi, vn, O.incident_field(ph, w, FP))[0]


def test_vtk_reader_cell_count():
    # mesh_data.rs:185-187: 336 cells in sphere.vtk; body 3 holds 324 triangles
    nodes, elems = O.read_vtk(SPHERE)
    assert len(elems) == 336
    assert len(O.body_elements(elems, 3)) == 324


def test_fixture_classical_lu(sphere_cpts):
    p = _fixture_field(sphere_cpts, O.CLASSICAL, O.SOLVER_DIRECT)
    assert abs(p.real - P_CLASSICAL.real) < 1e-16 and abs(p.imag - P_CLASSICAL.imag) < 1e-16
    assert abs(abs(p) - 0.0004776828778512934) < 1e-16
    assert abs(p - P_CLASSICAL) / abs(P_CLASSICAL) < 1e-12


def test_fixture_classical_gmres(sphere_cpts):
    p = _fixture_field(sphere_cpts, O.CLASSICAL, O.SOLVER_ITERATIVE)
    assert abs(p.real - P_CLASSICAL.real) < 1e-8 and abs(p.imag - P_CLASSICAL.imag) < 1e-8


def test_fixture_burton_miller_lu(sphere_cpts):
    p = _fixture_field(sphere_cpts, O.BURTON_MILLER, O.SOLVER_DIRECT)
    assert abs(p - P_BM) / abs(P_BM) < 1e-12
    assert abs(abs(p) - 0.0004964814282687255) < 1e-16


def test_fixture_hmatrix(sphere_cpts):
    p = _fixture_field(sphere_cpts, O.CLASSICAL, None, hier=True)
    assert abs(p.real - P_HMATRIX.real) < 1e-10 and abs(p.imag - P_HMATRIX.imag) < 1e-10
    assert abs(p - P_HMATRIX) / abs(P_HMATRIX) < 1e-12


def test_rng_lu_fixture():
    # solve.rs:211-218 (StdRng seed 10 -> 100x100 system -> x[0])
    a, b = O.generate_random_ab(100, 100, 10)
    x = O.lu_solve(a, b)
    assert x[0].real == pytest.approx(0.0035386336544537098, abs=1e-15)
    assert x[0].imag == pytest.approx(0.0017794290945383035, abs=1e-15)
    # and LAPACK agrees
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-12, atol=1e-14)


def test_rng_gmres_fixture():
    # solve.rs:220-229
    a, b = O.generate_random_ab(100, 100, 10)
    x, info = O.gmres(a, b, 100, 1e-8)
    assert abs(x[0].real - 0.0035386336544537098) < 1e-6
    assert abs(x[0].imag - 0.0017794290945383035) < 1e-6


def test_gmres_textbook_matches_lu():
    rng = np.random.default_rng(3)
    n = 200
    a = np.eye(n) * 8 + rng.standard_normal((n, n)) * 0.3 + 1j * rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x_ref = np.linalg.solve(a, b)
    for ref_mgs in (True, False):
        x, info = O.gmres(a, b, 1000, 1e-13, ref_mgs=ref_mgs)
        assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-10
    # the reported backward error is ||b - Ax|| / (||A||_F ||x|| + ||b||) (gmres.rs:142-144)
    for ref_mgs in (True, False):
        x, info = O.gmres(a, b, 1000, 1e-6, ref_mgs=ref_mgs)
        be = np.linalg.norm(b - a @ x) / (np.linalg.norm(a) * np.linalg.norm(x) + np.linalg.norm(b))
        assert info["backward_error"] == pytest.approx(be, rel=1e-6)
        assert be < 1e-6
    # the reference coefficient conj(q^H w) needs more matvecs than textbook MGS
    _, i_ref = O.gmres(a, b, 1000, 1e-10, ref_mgs=True)
    _, i_txt = O.gmres(a, b, 1000, 1e-10, ref_mgs=False)
    assert i_txt["num_mv"] <= i_ref["num_mv"]


def _grid(n, scale=1.0):
    pts = [[j * scale, k * scale, 0.0] for j in range(n) for k in range(n)]
    return np.array(pts)


def test_cluster_diameter():
    # cluster.rs:120-139: 25x25 integer grid -> root diameter 24*sqrt(2)
    _, diam = O.build_blocks(_grid(25), 32, 4.0)
    assert diam == pytest.approx(33.94, abs=1e-2)
    assert diam == pytest.approx(24 * math.sqrt(2), rel=1e-15)


def test_block_tree_admissibility():
    g = _grid(25)
    # block.rs:104-123: children[1][1][1] admissible
    assert O.block_tree_admissible_at(g, 32, 4.0, [1, 1, 1]) == 1
    # block.rs:125-145: flattened list[5] admissible
    blocks, _ = O.build_blocks(g, 32, 4.0)
    assert blocks[5][2] is True
    # every (row, col) pair is covered exactly once
    cover = np.zeros((625, 625), dtype=int)
    for r, c, _ in blocks:
        cover[np.ix_(r, c)] += 1
    assert (cover == 1).all()


def test_aca_rank13_fixture():
    # aca.rs:280-313: 20x100 rank-13 + 1e-8 noise
    m, n = 20, 100
    a, b = O.generate_random_ab(m, n, 13)
    a = a * 1e-8
    for k in range(13):
        _, c = O.generate_random_ab(m, m, k)
        _, d = O.generate_random_ab(n, n, k + 2)
        a = a + np.outer(c, d)
    res = O.aca(1e-6, a)
    assert res["norm"] == pytest.approx(np.linalg.norm(a), rel=0.05)
    x1 = a @ b
    x2 = res["U"] @ (res["V"].T @ b)
    assert np.linalg.norm(x2 - x1) / np.linalg.norm(x1) < 1e-8


def test_aca_vs_truncated_svd():
    # a smooth kernel block: ACA error must be of the order of the SVD tail
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (80, 3))
    y = rng.uniform(0, 1, (60, 3)) + np.array([4.0, 0, 0])
    d = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)
    A = np.exp(1j * 2.0 * d) / d
    res = O.aca(1e-6, A)
    err = np.linalg.norm(res["U"] @ res["V"].T - A) / np.linalg.norm(A)
    s = np.linalg.svd(A, compute_uv=False)
    r = res["rank"]
    tail = np.sqrt((s[r:] ** 2).sum()) / np.linalg.norm(s)
    assert err < 1e-5
    assert err < 100 * tail + 1e-12


def _hyperbolic(n):
    pts = np.array([[j / n, k / n, 0.0] for j in range(n) for k in range(n)])
    d = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(axis=2)
    return pts, (1.0 / np.maximum(1e-5, d)).astype(np.complex128)


def test_hmatrix_to_full_fixture():
    # h_matrix.rs:313-322: 10x10 grid, leaf 20, tol 1e-4 -> ||H - A|| / ||A|| < 1e-6
    pts, A = _hyperbolic(10)
    H = O.HMatrix(100, lambda r, c: A[np.ix_(r, c)].reshape(-1), pts, 20, 1e-4, 4.0)
    eps = np.linalg.norm(H.to_full() - A) / np.linalg.norm(A)
    assert eps < 1e-6


def test_hmatrix_mult_fixture():
    # h_matrix.rs:310-325: 20x20 grid matvec
    pts, A = _hyperbolic(20)
    _, b = O.generate_random_ab(400, 400, 10)
    H = O.HMatrix(400, lambda r, c: A[np.ix_(r, c)].reshape(-1), pts, 20, 1e-4, 4.0)
    x1 = A @ b
    x2 = H.gemv(1.0, b, 0.0, np.zeros(400, complex))
    assert np.linalg.norm(x2 - x1) / np.linalg.norm(x1) < 1e-6
    st = H.stats()
    assert st["admissible"] > 0 and 0 < st["compression_ratio"] < 100


def test_geometry_area_and_normals(sphere_cpts):
    cp = sphere_cpts
    # outward unit normals, total area close to 4 pi r^2 (faceted, r = 0.5)
    assert np.allclose(np.linalg.norm(cp[:, 3:6], axis=1), 1.0, atol=1e-14)
    assert (np.einsum("ij,ij->i", cp[:, 0:3], cp[:, 3:6]) > 0).all()
    area = (cp[:, 6] * cp[:, 7]).sum()
    assert 0.97 * math.pi < area < math.pi


def test_quad_quarter_area_quirk():
    # elements.rs:168: quads integrate to a quarter of their area
    cp = O.mesh_cpts_from_vtk(os.path.join(GOLD, "sphere_quad.vtk"), 2)
    assert cp.shape[0] == 255
    assert np.all(cp[:, 7] == 4.0)
    wa = (cp[:, 6] * cp[:, 7]).sum()
    assert 0.74 < wa < 0.79           # ~ (pi r^2 * 4) / 4 with r = 0.5


def test_kernel_reciprocity_and_gauss(sphere_cpts):
    cp = sphere_cpts
    n = cp.shape[0]
    ph = O.physics(1e-3, 343.0, 1.21)   # k -> 0
    H, G = O.dense_surface_matrices(ph, cp)
    wa = cp[:, 6] * cp[:, 7]
    Gk = G / wa[None, :]
    off = ~np.eye(n, dtype=bool)
    np.testing.assert_allclose(Gk[off], Gk.T[off], rtol=1e-13)
    # Gauss integral: sum_j != i H_ij ~ -1/2 at k -> 0 for a closed surface
    s = (H * off).sum(axis=1)
    assert abs(s.real.mean() + 0.5) < 0.05


def test_pulsating_sphere_closed_form():
    # classical, unit normal velocity on r = 0.5 sphere (refined mesh, n = 1120)
    cp = O.mesh_cpts_from_vtk(os.path.join(GOLD, "refined_sphere.vtk"), 3)
    a = 0.5
    c = 343.0
    ka = 1.0
    f = ka / a * c / (2 * math.pi)
    ph = O.physics(f, c, 1.21, O.CLASSICAL, O.EXTERIOR, O.BC_NORMAL_VELOCITY, (1.0, 0.0))
    rhs = np.zeros(cp.shape[0], complex)
    phi, vn, _ = O.solve_surface_dense(ph, cp, rhs, O.SOLVER_DIRECT)
    exact = 1.0 * a / (1j * ka - 1.0)
    err = np.abs(phi.mean() - exact) / abs(exact)
    assert err < 0.03
    assert np.allclose(vn, 1.0)


def test_bc_algebra_consistency(sphere_cpts):
    """Pressure, velocity and impedance branches satisfy their own equations."""
    cp = sphere_cpts
    w = O.wave(O.WAVE_PLANE, (0.3, 0.4, 0.5), (1.0, 0.5))
    for bc, val in ((O.BC_PRESSURE, (0.2, -0.1)), (O.BC_NORMAL_VELOCITY, (0.5, 0.2)),
                    (O.BC_IMPEDANCE, (415.0, 83.0))):
        ph = O.physics(200.0, 343.0, 1.21, O.BURTON_MILLER, O.EXTERIOR, bc, val)
        rhs = O.incident_surface(ph, w, cp)
        phi, vn, _ = O.solve_surface_dense(ph, cp, rhs, O.SOLVER_DIRECT)
        H, G = O.dense_surface_matrices(ph, cp)
        # each branch satisfies its own equation of solve.rs:36-69
        if bc == O.BC_IMPEDANCE:
            omega = 2 * math.pi * 200.0
            f = (-1j * omega * 1.21) / complex(*val)
            np.testing.assert_allclose((H + f * G) @ phi, -rhs, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(vn, f * phi, rtol=1e-14)
        elif bc == O.BC_NORMAL_VELOCITY:   # H phi = G vn - rhs_inc
            assert np.allclose(vn, complex(*val))
            res = H @ phi - G @ vn + rhs
            assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(rhs)
        else:                              # G vn = rhs_inc - H phi
            assert np.allclose(phi, complex(*val))
            res = G @ vn + H @ phi - rhs
            assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(rhs)


# --------------------------------------------------------------------------
# Closed forms for the impedance branch (solve.rs:57-68), the spherical
# incident wave (incident_wave.rs:40-62) and the pressure branch
# (solve.rs:36-43).  The series in oracle/analytic.py are pinned first by
# the conditions that define them, then the oracle is held against them
# within the discretisation error of the 1120-triangle sphere (about 2 %).
# --------------------------------------------------------------------------
from oracle import analytic as AN  # noqa: E402

REFINED = os.path.join(GOLD, "refined_sphere.vtk")   # radius 0.5, 1120 triangles
RHO, C0, A_R = 1.21, 343.0, 0.5


@pytest.fixture(scope="module")
def refined_cpts():
    return O.mesh_cpts_from_vtk(REFINED, 3)


def _on_sphere(cp):
    """Collocation points projected radially onto r = a (the centroids of the
    faceted sphere sit 0.4 % inside it)."""
    x = cp[:, :3]
    return A_R * x / np.linalg.norm(x, axis=1)[:, None]


def test_analytic_impedance_series_is_pinned():
    k, sigma = 2.0, 0.3846 + 1.923j
    ct_ = np.linspace(-1, 1, 7)
    hh = 1e-5
    # d(phi)/dr = sigma phi at r = a (central difference)
    d = (AN.impedance_sphere_total(A_R, k, sigma, A_R + hh, ct_)
         - AN.impedance_sphere_total(A_R, k, sigma, A_R - hh, ct_)) / (2 * hh)
    phi = AN.impedance_sphere_total(A_R, k, sigma, A_R, ct_)
    np.testing.assert_allclose(d, sigma * phi, rtol=1e-7, atol=1e-9)
    # sigma = 0 is the rigid sphere: incident minus the reference-signed
    # scattered series of the module's first function
    r = 0.9
    rigid = np.exp(1j * k * r * ct_) - AN.rigid_sphere_scattered(A_R, k, r, ct_)
    np.testing.assert_allclose(AN.impedance_sphere_total(A_R, k, 0.0, r, ct_), rigid,
                               rtol=1e-12, atol=1e-12)
    # the truncated series has converged
    np.testing.assert_allclose(AN.impedance_sphere_total(A_R, k, sigma, r, ct_, nterms=40),
                               AN.impedance_sphere_total(A_R, k, sigma, r, ct_, nterms=60),
                               rtol=1e-13, atol=1e-13)


def test_analytic_point_source_series_is_pinned():
    k, rs = 3.0, 1.0
    p = np.array([[0.5, 0.1, -0.2], [0.0, 0.45, 0.2], [-0.3, -0.3, 0.3]])
    s = np.array([0.0, 0.0, rs])
    R = np.linalg.norm(p - s, axis=1)
    r = np.linalg.norm(p, axis=1)
    direct = np.exp(1j * k * R) / (4 * np.pi * R)
    np.testing.assert_allclose(AN.point_source_series(k, rs, r, p[:, 2] / r), direct,
                               rtol=1e-10)
    # rigid surface: the radial derivative of the total field vanishes at r = a
    u = p / r[:, None]
    hh = 1e-5
    dr = (AN.point_source_rigid_sphere_total(A_R, k, s, (A_R + hh) * u)
          - AN.point_source_rigid_sphere_total(A_R, k, s, (A_R - hh) * u)) / (2 * hh)
    scale = np.abs(AN.point_source_rigid_sphere_total(A_R, k, s, A_R * u)).max() * k
    assert np.abs(dr).max() < 1e-7 * scale


@pytest.mark.parametrize("ka", [0.5, 1.0, 2.0])
def test_impedance_branch_vs_impedance_sphere(refined_cpts, ka):
    """Classical formulation, impedance Z = rho c (1 + 0.2i), plane wave:
    the oracle's phi against the impedance-sphere series with
    sigma = i omega rho / Z; the opposite sign of the f G term is 60-130 %
    off, so a sign slip in solve.rs:57-68's restatement fails here."""
    cp = refined_cpts
    k = ka / A_R
    f = k * C0 / (2 * math.pi)
    omega = 2 * math.pi * f
    Z = complex(RHO * C0, 0.2 * RHO * C0)
    sigma = 1j * omega * RHO / Z
    d = np.array([0.3, 0.4, 0.5])
    ph = O.physics(f, C0, RHO, O.CLASSICAL, O.EXTERIOR, O.BC_IMPEDANCE, (Z.real, Z.imag))
    w = O.wave(O.WAVE_PLANE, tuple(d), (1.0, 0.0))
    phi, vn, _ = O.solve_surface_dense(ph, cp, O.incident_surface(ph, w, cp))
    amp = 1.0 / (1j * omega * RHO)
    x = _on_sphere(cp)
    exact = AN.impedance_sphere_total_at(x, A_R, k, sigma, d, amp)
    err = np.linalg.norm(phi - exact) / np.linalg.norm(exact)
    assert err < 0.03, err
    wrong = AN.impedance_sphere_total_at(x, A_R, k, -sigma, d, amp)
    assert np.linalg.norm(phi - wrong) / np.linalg.norm(wrong) > 0.5
    # the returned vn is f phi = -sigma phi: the opposite of the normal
    # velocity the solved system implies (the reference's quirk, solve.rs:67)
    np.testing.assert_allclose(vn, -sigma * phi, rtol=1e-13)
    assert np.linalg.norm(vn + sigma * exact) / np.linalg.norm(sigma * exact) < 0.03


@pytest.mark.parametrize("ka", [0.5, 1.0, 2.0])
def test_spherical_wave_vs_point_source_series(refined_cpts, ka):
    """Classical, rigid sphere of radius 0.5 in the field of the spherical
    wave: the oracle moves the source to origin/|origin| (incident_wave.rs:44),
    here a point outside the body, so the total surface potential is the
    point-source rigid-sphere series."""
    cp = refined_cpts
    k = ka / A_R
    f = k * C0 / (2 * math.pi)
    omega = 2 * math.pi * f
    origin = np.array([0.2, -0.6, 0.5])
    ph = O.physics(f, C0, RHO, O.CLASSICAL, O.EXTERIOR, O.BC_NORMAL_VELOCITY, (0.0, 0.0))
    w = O.wave(O.WAVE_SPHERICAL, tuple(origin), (0.7, -0.3))
    phi, _, _ = O.solve_surface_dense(ph, cp, O.incident_surface(ph, w, cp))
    amp = complex(0.7, -0.3) / (1j * omega * RHO)
    exact = AN.point_source_rigid_sphere_total(A_R, k, origin / np.linalg.norm(origin),
                                               _on_sphere(cp), amp)
    err = np.linalg.norm(phi - exact) / np.linalg.norm(exact)
    assert err < 0.03, err


@pytest.mark.parametrize("origin", [(0.2, -0.6, 0.5), (0.05, 0.01, -0.02)])
def test_spherical_wave_bm_term_finite_difference(refined_cpts, origin):
    """Burton-Miller right-hand side of the spherical wave
    (incident_wave.rs:53-58): (rhs_BM - rhs_classical) / beta, beta = i/k,
    equals the normal derivative of the classical phi_I at each collocation
    point, taken by central differences along the normal."""
    cp = refined_cpts
    f = 300.0
    k = 2 * math.pi * f / C0
    w = O.wave(O.WAVE_SPHERICAL, origin, (0.7, -0.3))
    phC = O.physics(f, C0, RHO, O.CLASSICAL)
    phB = O.physics(f, C0, RHO, O.BURTON_MILLER)
    term = (O.incident_surface(phB, w, cp) - O.incident_surface(phC, w, cp)) / (1j / k)
    hh = 1e-5
    cpp, cpm = cp.copy(), cp.copy()
    cpp[:, :3] += hh * cp[:, 3:6]
    cpm[:, :3] -= hh * cp[:, 3:6]
    fd = (O.incident_surface(phC, w, cpp) - O.incident_surface(phC, w, cpm)) / (2 * hh)
    assert np.linalg.norm(term - fd) / np.linalg.norm(fd) < 1e-7


def test_plane_wave_bm_term_finite_difference(refined_cpts):
    """The same check for the plane wave's Burton-Miller term
    (incident_wave.rs:30-36)."""
    cp = refined_cpts
    f = 300.0
    k = 2 * math.pi * f / C0
    w = O.wave(O.WAVE_PLANE, (0.3, -0.4, 0.5), (0.7, -0.3))
    phC = O.physics(f, C0, RHO, O.CLASSICAL)
    phB = O.physics(f, C0, RHO, O.BURTON_MILLER)
    term = (O.incident_surface(phB, w, cp) - O.incident_surface(phC, w, cp)) / (1j / k)
    hh = 1e-5
    cpp, cpm = cp.copy(), cp.copy()
    cpp[:, :3] += hh * cp[:, 3:6]
    cpm[:, :3] -= hh * cp[:, 3:6]
    fd = (O.incident_surface(phC, w, cpp) - O.incident_surface(phC, w, cpm)) / (2 * hh)
    assert np.linalg.norm(term - fd) / np.linalg.norm(fd) < 1e-7


@pytest.mark.parametrize("ka", [0.5, 1.0])
def test_pressure_branch_sign_quirk(refined_cpts, ka):
    """Dirichlet pulsating sphere: phi = p_bar on r = a, outgoing field, so
    the physical normal velocity is d(phi)/dr = p_bar (ik - 1/a).  The
    reference's pressure branch solves G vn = rhs_inc - H p_bar
    (solve.rs:36-43) where its velocity branch (solve.rs:46-52) implies
    G vn = rhs_inc + H p_bar, so its vn comes out with the opposite sign:
    mean vn = -p_bar (ik - 1/a) within the discretisation error."""
    cp = refined_cpts
    k = ka / A_R
    f = k * C0 / (2 * math.pi)
    pbar = complex(0.8, -0.3)
    ph = O.physics(f, C0, RHO, O.CLASSICAL, O.EXTERIOR, O.BC_PRESSURE, (pbar.real, pbar.imag))
    phi, vn, _ = O.solve_surface_dense(ph, cp, np.zeros(cp.shape[0], complex))
    physical = pbar * (1j * k - 1.0 / A_R)
    assert np.allclose(phi, pbar)
    # (pointwise vn scatters by 20 % about its mean: G, the single-layer
    # operator with one-point quadrature, is a first-kind equation)
    assert abs(vn.mean() + physical) / abs(physical) < 0.02
    assert abs(vn.mean() - physical) / abs(physical) > 1.9
