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

Two more textbook series pin the oracle's other branches:

* ``impedance_sphere_total``: a plane wave on a sphere whose surface obeys
  d(phi)/dr = sigma phi (a locally reacting surface).  The reference's
  impedance system (H + f G) phi = -rhs_inc, f = -i omega rho / Z
  (solve.rs:57-68) is the velocity equation H phi = G v - rhs_inc
  (solve.rs:46-52) with v = -f phi, i.e. sigma = i omega rho / Z;
* ``point_source_rigid_sphere_total``: a rigid sphere in the field of a point
  source outside it, phi_I = A e^{ikR}/(4 pi R) (incident_wave.rs:40-52).

Both are total fields (incident + scattered) with the time factor
e^{-i omega t}, which is the unknown the reference's velocity / impedance
systems solve for.  Only tests/ use this module.
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


def impedance_sphere_total(a: float, k: float, sigma: complex, r, cos_theta,
                           amplitude=1.0 + 0j, nterms: int = 60):
    """Total potential phi_I + phi_s at radius r >= a for the plane wave
    amplitude * e^{i k r cos(theta)} on a sphere with d(phi)/dr = sigma phi
    at r = a:

        phi = A sum (2n+1) i^n P_n [j_n(kr) + c_n h_n(kr)],
        c_n = -(k j_n'(ka) - sigma j_n(ka)) / (k h_n'(ka) - sigma h_n(ka)).

    sigma = 0 is the rigid sphere."""
    r = np.asarray(r, dtype=np.float64)
    ct_ = np.asarray(cos_theta, dtype=np.float64)
    out = np.zeros(np.broadcast(r, ct_).shape, dtype=np.complex128)
    ka = k * a
    for n in range(nterms):
        cn = -(k * spherical_jn(n, ka, derivative=True) - sigma * spherical_jn(n, ka)) / \
            (k * _h1(n, ka, True) - sigma * _h1(n, ka))
        radial = spherical_jn(n, k * r) + cn * _h1(n, k * r)
        out += (2 * n + 1) * (1j ** n) * eval_legendre(n, ct_) * radial
    return amplitude * out


def impedance_sphere_total_at(points, a: float, k: float, sigma: complex, direction,
                              amplitude=1.0 + 0j, nterms: int = 60):
    """``impedance_sphere_total`` at Cartesian points (sphere at the origin)."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    r = np.linalg.norm(p, axis=1)
    return impedance_sphere_total(a, k, sigma, r, (p @ d) / r, amplitude, nterms)


def point_source_series(k: float, rs: float, r, cos_gamma, amplitude=1.0 + 0j,
                        nterms: int = 80):
    """A e^{ikR}/(4 pi R) for r < rs by the addition theorem,
    (i k A / 4 pi) sum (2n+1) j_n(kr) h_n(k rs) P_n(cos gamma); tests compare
    it with the closed form to pin the expansion used below."""
    r = np.asarray(r, dtype=np.float64)
    ct_ = np.asarray(cos_gamma, dtype=np.float64)
    out = np.zeros(np.broadcast(r, ct_).shape, dtype=np.complex128)
    for n in range(nterms):
        out += (2 * n + 1) * spherical_jn(n, k * r) * _h1(n, k * rs) * eval_legendre(n, ct_)
    return amplitude * 1j * k / (4 * np.pi) * out


def point_source_rigid_sphere_total(a: float, k: float, source, points,
                                    amplitude=1.0 + 0j, nterms: int = 80):
    """Total potential at points (a <= |x| < |source|) around a rigid sphere
    of radius a at the origin, incident field A e^{ikR}/(4 pi R) from
    `source` (|source| > a):

        phi = (i k A / 4 pi) sum (2n+1) h_n(k rs) P_n(cos gamma)
              [j_n(kr) - j_n'(ka) h_n(kr) / h_n'(ka)]."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    s = np.asarray(source, dtype=np.float64)
    rs = np.linalg.norm(s)
    r = np.linalg.norm(p, axis=1)
    cg = (p @ s) / (r * rs)
    out = np.zeros(p.shape[0], dtype=np.complex128)
    ka = k * a
    for n in range(nterms):
        radial = spherical_jn(n, k * r) - spherical_jn(n, ka, derivative=True) / \
            _h1(n, ka, True) * _h1(n, k * r)
        out += (2 * n + 1) * _h1(n, k * rs) * eval_legendre(n, cg) * radial
    return amplitude * 1j * k / (4 * np.pi) * out
