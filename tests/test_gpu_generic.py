"""The closure-generic ACA and H matrix of the reference over explicit
matrices (aca.rs:22-33 ACA::new, h_matrix.rs:127-148 HMatrix::new_from),
replaying the reference's own tests through the C ABI:

* aca.rs:280-313 (mult_aca): a 20 x 100 rank-13 matrix plus 1e-8 noise built
  from rand 0.9's StdRng (the oracle's bit-exact restatement generates it),
  ACA at 1e-6: norm estimate within 5 % of ||A||_F, matvec within 1e-8;
* h_matrix.rs:300-324 (get_hmatrix, mult_hmatrix): 1/d^2 on 10 x 10 and
  20 x 20 grids, leaf 20, tolerance 1e-4: to_full and matvec within 1e-6.

Both are also held against the oracle's restatement of the same algorithm.
"""
import numpy as np
import pytest
import torch

import oracle as O
from cc_abstemious_amd import aca, capi, h_matrix, solve

pytestmark = pytest.mark.gpu


def _note(msg):
    print(f"[case] {msg}", flush=True)


def _rank13():
    m, n = 20, 100
    a, b = O.generate_random_ab(m, n, 13)
    a = a * 1e-8
    for k in range(13):
        _, c = O.generate_random_ab(m, m, k)
        _, d = O.generate_random_ab(n, n, k + 2)
        a = a + np.outer(c, d)
    return a, b


def _hyperbolic(n):
    pts = np.array([[j / n, k / n, 0.0] for j in range(n) for k in range(n)])
    d = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(axis=2)
    return pts, (1.0 / np.maximum(1e-5, d)).astype(np.complex128)


def test_aca_reference_fixture():
    _note("ACA aca.rs:280-313 rank-13 20x100, tol 1e-6")
    a, b = _rank13()
    A = torch.from_numpy(a).cuda()
    ac = aca.ACA(1e-6, A)
    st = ac.stats()
    assert st["dense_fallback"] == 0 and st["admissible"] == 1 and st["inadmissible"] == 0
    norm = ac.get_norm()
    assert norm == pytest.approx(np.linalg.norm(a), rel=0.05)      # aca.rs:306
    x1 = a @ b
    x2 = torch.zeros(20, dtype=torch.complex128, device="cuda")
    ac.gemv(1.0, torch.from_numpy(b).cuda(), 0.0, x2)
    f = np.linalg.norm(x2.cpu().numpy() - x1) / np.linalg.norm(x1)
    print(f"rank {ac.get_num_uv()}, norm {norm:.6e} vs {np.linalg.norm(a):.6e}, matvec {f:.2e}")
    assert f < 1e-8                                                 # aca.rs:312
    # the oracle takes the same pivots: same rank, norm and factors' product
    ref = O.aca(1e-6, a)
    assert ac.get_num_uv() == ref["rank"]
    assert norm == pytest.approx(ref["norm"], rel=1e-10)
    full = ac.to_full().cpu().numpy()
    uv = ref["U"] @ ref["V"].T
    assert np.linalg.norm(full - uv) / np.linalg.norm(uv) < 1e-12
    # beta and alpha of the matvec (aca.rs:181-195 scales b by beta first)
    y0 = torch.from_numpy(np.arange(20) * (1 + 1j)).cuda()
    y = y0.clone()
    ac.gemv(2.0 - 1.0j, torch.from_numpy(b).cuda(), 0.5j, y)
    want = (2.0 - 1.0j) * (uv @ b) + 0.5j * y0.cpu().numpy()
    np.testing.assert_allclose(y.cpu().numpy(), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_aca_rectangular_tall_and_smooth():
    """A tall smooth kernel block (80 x 60 source / target clouds apart): the
    ACA error is of the order of the truncated SVD's tail, as for the
    oracle."""
    _note("ACA smooth 80x60 and tall 200x30")
    rng = np.random.default_rng(0)
    for m, n in ((80, 60), (200, 30)):
        x = rng.uniform(0, 1, (m, 3))
        y = rng.uniform(0, 1, (n, 3)) + np.array([4.0, 0, 0])
        d = np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)
        a = np.exp(1j * 2.0 * d) / d
        ac = aca.ACA(1e-6, torch.from_numpy(a).cuda())
        full = ac.to_full().cpu().numpy()
        assert full.shape == (m, n)
        err = np.linalg.norm(full - a) / np.linalg.norm(a)
        ref = O.aca(1e-6, a)
        assert ac.get_num_uv() == ref["rank"]
        assert err < 1e-5
        np.testing.assert_allclose(full, ref["U"] @ ref["V"].T, rtol=0,
                                   atol=1e-11 * np.abs(a).max())


@pytest.mark.parametrize("grid,check", [(10, "to_full"), (20, "matvec")])
def test_hmatrix_reference_fixture(grid, check):
    """h_matrix.rs:300-324: leaf 20, tolerance 1e-4 (and the fixed eta 4)."""
    _note(f"H matrix h_matrix.rs {check} fixture, {grid}x{grid} grid")
    pts, a = _hyperbolic(grid)
    n = grid * grid
    A = torch.from_numpy(a).cuda()
    H = h_matrix.HMatrix.new_from_matrix(A, pts, 20, 1e-4, 4.0)
    Ho = O.HMatrix(n, lambda r, c: a[np.ix_(r, c)].reshape(-1), pts, 20, 1e-4, 4.0)
    st, sto = H.stats(), Ho.stats()
    print("gpu", st, "oracle", sto)
    assert st["dense_fallback"] == 0
    assert (st["admissible"], st["inadmissible"]) == (sto["admissible"], sto["inadmissible"])
    full = H.to_full().cpu().numpy()
    if check == "to_full":
        eps = np.linalg.norm(full - a) / np.linalg.norm(a)
        print(f"to_full vs A: {eps:.2e}")
        assert eps < 1e-6                                            # h_matrix.rs:308
    else:
        _, b = O.generate_random_ab(n, n, 10)
        x1 = a @ b
        x2 = torch.zeros(n, dtype=torch.complex128, device="cuda")
        H.gemv(1.0, torch.from_numpy(b).cuda(), 0.0, x2)
        f = np.linalg.norm(x2.cpu().numpy() - x1) / np.linalg.norm(x1)
        print(f"matvec vs A b: {f:.2e}")
        assert f < 1e-6                                              # h_matrix.rs:323
    # the oracle's H matrix represents the same matrix to the ACA tolerance
    # (grid points tie exactly in the pivot searches, so the pivots may differ)
    e2 = np.linalg.norm(full - Ho.to_full()) / np.linalg.norm(a)
    assert e2 < 1e-6


def test_gmres_over_generic_hmatrix():
    """GMRES (gmres.rs:35-202) with a generic H matrix as the operator
    (gm.hmatrix = Some(..), solve.rs:119-125) against numpy's LAPACK solve
    of the dense matrix."""
    _note("GMRES over the generic H matrix (20x20 grid)")
    pts, a = _hyperbolic(20)
    n = 400
    H = h_matrix.HMatrix.new_from_matrix(torch.from_numpy(a).cuda(), pts, 20, 1e-8, 4.0)
    _, b = O.generate_random_ab(n, n, 10)
    gm = solve.GMRES(400, 1e-12)
    gm.hmatrix = H
    x = torch.from_numpy(b.copy()).cuda()
    gm.solve(x)
    want = np.linalg.solve(a, b)
    err = np.linalg.norm(x.cpu().numpy() - want) / np.linalg.norm(want)
    print(f"GMRES over H vs LAPACK: {err:.2e}")
    assert err < 1e-6


def test_generic_boundary_errors():
    ctx = capi.default_context(0)
    A = torch.zeros((4, 3), dtype=torch.complex128, device="cuda")
    h = __import__("ctypes").c_void_p()
    with pytest.raises(capi.CcabError):   # lda < m
        ctx.call("ccab_aca_build", 4, 3, h_matrix._ptr(A), 2, 1e-6, __import__("ctypes").byref(h))
    with pytest.raises(capi.CcabError):   # tolerance must be positive
        ctx.call("ccab_aca_build", 4, 3, h_matrix._ptr(A), 4, 0.0, __import__("ctypes").byref(h))
    # GMRES needs a square operator
    a = np.random.default_rng(1).standard_normal((6, 4)) + 0j
    ac = aca.ACA(1e-6, torch.from_numpy(a).cuda())
    gm = solve.GMRES(10, 1e-8)
    gm.hmatrix = ac
    with pytest.raises((capi.CcabError, ValueError)):
        gm.solve(torch.ones(6, dtype=torch.complex128, device="cuda"))
