"""Pins the FP64 surrogate oracle (oracle/surrogate_ref.py) on the CPU before
it checks the HIP training kernels: the analytic gradient against central
finite differences, BF16 rounding against its definition, Adam against its
closed form on a 1-parameter problem, and the feature map against geometric
identities."""
import numpy as np
import pytest

import oracle as O
from oracle import surrogate_ref as S


def _net(W, seed):
    rng = np.random.default_rng(seed)
    offs, total = S.layout(W)
    p = np.zeros(total)
    for ow, ob, fo, fi in offs:
        p[ow:ow + fo * fi] = rng.uniform(-1, 1, fo * fi) / np.sqrt(fi)
        p[ob:ob + fo] = rng.uniform(-0.1, 0.1, fo)
    return p


def test_gradient_matches_finite_differences():
    W, P = 64, 37
    rng = np.random.default_rng(1)
    p = _net(W, 2)
    X = rng.standard_normal((P, S.FEAT))
    T = rng.standard_normal((P, 2))
    loss, g = S.loss_grad(p, X, T, W, emulate=False)
    assert abs(loss - S.loss_only(p, X, T, W)) < 1e-14 * max(1.0, loss)
    offs, total = S.layout(W)
    # a sample of entries from every weight and bias block
    idx = []
    for ow, ob, fo, fi in offs:
        idx += list(rng.choice(np.arange(ow, ow + fo * fi), 6, replace=False))
        idx += list(rng.choice(np.arange(ob, ob + fo), 2, replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros(total)
        e[i] = h
        fd = (S.loss_only(p + e, X, T, W) - S.loss_only(p - e, X, T, W)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_gradient_directional_derivative():
    W, P = 64, 50
    rng = np.random.default_rng(3)
    p = _net(W, 4)
    X = rng.standard_normal((P, S.FEAT))
    T = rng.standard_normal((P, 2))
    _, g = S.loss_grad(p, X, T, W)
    d = rng.standard_normal(p.size)
    h = 1e-6
    fd = (S.loss_only(p + h * d, X, T, W) - S.loss_only(p - h * d, X, T, W)) / (2 * h)
    assert abs(fd - g @ d) <= 1e-7 * abs(fd)


def test_bf16_rounding():
    one = 1.0
    ulp = 2.0 ** -7                      # bf16 spacing in [1, 2)
    assert S.bf16(one + ulp / 2) == one              # tie -> even (1.0)
    assert S.bf16(one + 1.5 * ulp) == one + 2 * ulp  # tie -> even (1 + 2 ulp)
    assert S.bf16(one + 0.49 * ulp) == one
    assert S.bf16(one + 0.51 * ulp) == one + ulp
    x = np.random.default_rng(0).standard_normal(1000)
    r = S.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-7))
    assert np.all(S.bf16(r) == r)


def test_adam_closed_form_first_step():
    # first Adam step moves every parameter by lr * sign(g) (bias-corrected)
    W = 64
    p = _net(W, 5)
    g = np.random.default_rng(6).standard_normal(p.size)
    p1, m, v = S.adam_step(p, np.zeros_like(p), np.zeros_like(p), g, 1, W, lr=1e-3, eps=0.0)
    np.testing.assert_allclose(p1 - p, -1e-3 * np.sign(g), rtol=1e-12, atol=0)


def test_features_geometry_identities():
    cp = O.mesh_cpts_from_vtk("tests/golden/sphere.vtk", 3)
    sp = S.system_params(cp, 100.0, 1500.0, 1000.0, (1.0, 2.0, 2.0), (1.0, 0.0))
    f = S.features(cp, sp)
    np.testing.assert_allclose(np.linalg.norm(f[:, 6:9], axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(f[:, 2], f[:, 6:9] @ (np.array([1, 2, 2]) / 3.0), atol=1e-12)
    np.testing.assert_allclose(f[:, 13] ** 2 + f[:, 14] ** 2, 1.0, atol=1e-12)
    np.testing.assert_allclose(np.cos(np.pi * f[:, 12]), f[:, 2], atol=1e-12)
    # RMS radius: mean of (1 + f5)^2 is 1 by construction
    assert abs(np.mean((1 + f[:, 5]) ** 2) - 1.0) < 1e-12
    # the target of the incident field itself is psi = 1
    t = S.targets(cp, sp, S.incident(cp, sp))
    np.testing.assert_allclose(t, np.tile([1.0, 0.0], (cp.shape[0], 1)), atol=1e-14)
    # normalised targets: unit RMS and invariant to the solution's scale
    phi = S.incident(cp, sp) * (1.0 + 0.3 * np.cos(3 * f[:, 12]))
    t1, t2 = S.targets(cp, sp, phi), S.targets(cp, sp, (2.0 - 1.5j) * phi)
    assert abs(np.mean(np.sum(t1 ** 2, axis=1)) - 1.0) < 1e-12
    np.testing.assert_allclose(np.abs(t1[:, 0] + 1j * t1[:, 1]), np.abs(t2[:, 0] + 1j * t2[:, 1]),
                               rtol=1e-12)
    # area feature: relative element area
    np.testing.assert_allclose(f[:, 26], cp[:, 6] / cp[:, 6].mean() - 1.0, atol=1e-15)
