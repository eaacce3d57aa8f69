"""Host checks of the opt-in corrections: the quad-area fix of the
collocation points (ccab_collocation_points_ex) against exact element areas,
and the physics struct field that carries the correction bits."""
import ctypes as ct

import numpy as np

from cc_abstemious_amd import capi, mesh


def _flat_parallelograms():
    """A 3 x 2 grid of sheared parallelograms in the plane z = 0.3x + 0.1y."""
    a = np.array([0.7, 0.1, 0.0])
    b = np.array([0.25, 0.5, 0.0])
    nodes = []
    for j in range(3):
        for i in range(4):
            p = i * a + j * b
            nodes.append([p[0], p[1], 0.3 * p[0] + 0.1 * p[1]])
    nodes = np.array(nodes)
    conn, nv = [], []
    for j in range(2):
        for i in range(3):
            q = [j * 4 + i, j * 4 + i + 1, (j + 1) * 4 + i + 1, (j + 1) * 4 + i]
            conn.append(q)
            nv.append(4)
    return mesh.Mesh(nodes, np.array(conn), np.array(nv, dtype=np.int32))


def test_quad_area_exact_on_parallelograms():
    m = _flat_parallelograms()
    exact = []
    for q in m.conn:
        p0, p1, p3 = m.nodes[q[0]], m.nodes[q[1]], m.nodes[q[3]]
        exact.append(np.linalg.norm(np.cross(p1 - p0, p3 - p0)))
    exact = np.array(exact)
    fixed = m.collocation_points(capi.FIX_QUAD_AREA)
    parity = m.collocation_points(0)
    np.testing.assert_allclose(fixed[:, 6] * fixed[:, 7], exact, rtol=1e-14)
    np.testing.assert_array_equal(fixed[:, 7], 1.0)
    # the reference's QUADGP1 (elements.rs:168): wt 4, area a sixteenth,
    # so wt * area is a quarter of the element's area
    np.testing.assert_allclose(parity[:, 6] * parity[:, 7], exact / 4, rtol=1e-14)
    # centroids and normals do not change
    np.testing.assert_array_equal(fixed[:, 0:6], parity[:, 0:6])


def test_quad_area_sphere_total():
    """Full quad areas of a cube-sphere sum to the sphere's area as the mesh
    refines (inscribed, so from below); triangles ignore the flag."""
    tot = []
    for mm in (8, 16, 32):
        c = mesh.cube_sphere(mm).collocation_points(capi.FIX_QUAD_AREA)
        tot.append(float((c[:, 6] * c[:, 7]).sum()))
    err = [abs(t - 4 * np.pi) / (4 * np.pi) for t in tot]
    assert err[0] > err[1] > err[2] and err[2] < 2e-3
    assert all(t < 4 * np.pi for t in tot)
    t = mesh.icosphere(2)
    np.testing.assert_array_equal(t.collocation_points(capi.FIX_QUAD_AREA), t.collocation_points())


def test_physics_struct_carries_corrections():
    assert ct.sizeof(capi.Physics) == 56
    ph = capi.physics(100.0, corrections=capi.FIX_ALL)
    assert ph.corrections == 63
    assert capi.Physics.corrections.offset == 36
