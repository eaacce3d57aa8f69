"""Closed-form rigid-sphere plane-wave scattering (test infrastructure only).

The reference's verification note (docs/verification/rigid_sphere_plane_wave/
rigid_sphere_plane_wave.md) compares its BEM field with the series

    p_s(r, theta) = p_i sum_n (2n+1) i^n P_n(cos theta) j_n'(ka) / h_n'(ka) h_n(kr)

(h_n = h_n^(1), time factor e^{-i omega t}, theta measured from the direction
of propagation).  Its sign is the opposite of the textbook scattered field,
-p_i sum(...), and matches the reference's field evaluation, whose
-M phi + L vn carries the same opposite sign (solve.rs:173-185).  This module
restates that series with scipy's spherical Bessel functions; the tests pin it
by the rigid-surface condition d/dr (p_i e^{ikr cos theta} - p_s) = 0 at r = a
(the derivative of the textbook total field) and by its convergence in n.

Only tests/ use this module.
"""
from __future__ import annotations

import numpy as np
from scipy.special import eval_legendre, spherical_jn, spherical_yn


def _h1(n, z, derivative=False):
    return spherical_jn(n, z, derivative=derivative) + 1j * spherical_yn(n, z, derivative=derivative)


def rigid_sphere_scattered(a: float, k: float, r, cos_theta, amplitude=1.0 + 0j, nterms: int = 60):
    """p_s at radius r (> a) and angle theta from the propagation direction,
    with the reference's sign (see the module docstring)."""
    r = np.asarray(r, dtype=np.float64)
    ct_ = np.asarray(cos_theta, dtype=np.float64)
    out = np.zeros(np.broadcast(r, ct_).shape, dtype=np.complex128)
    ka = k * a
    for n in range(nterms):
        coef = (2 * n + 1) * (1j ** n) * spherical_jn(n, ka, derivative=True) / _h1(n, ka, True)
        out += coef * eval_legendre(n, ct_) * _h1(n, k * r)
    return amplitude * out


def rigid_sphere_scattered_at(points, a: float, k: float, direction, amplitude=1.0 + 0j,
                              center=(0.0, 0.0, 0.0), nterms: int = 60):
    """The same series at Cartesian points, for a plane wave along
    `direction` scattered by a sphere of radius a at `center`."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3) - np.asarray(center, dtype=np.float64)
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    r = np.linalg.norm(p, axis=1)
    return rigid_sphere_scattered(a, k, r, (p @ d) / r, amplitude, nterms)
