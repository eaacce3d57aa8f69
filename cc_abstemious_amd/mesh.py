"""Surface meshes: the reference's ASCII-VTK reader and deterministic generators.

``read_vtk`` follows Mesh::read_from_vtk (src/preprocess/mesh_data.rs:96-175):
the first integer of each CELLS line is the body id (bodies numbered 1, 2, 3..
in order of first appearance), CELL_TYPES 5 = triangle, 9 = quad.
``write_vtk`` writes a file that reader accepts (dummy point/line cells fill
bodies 1 and 2 so a surface mesh can live in body 3, as the Gmsh meshes of the
reference's tests do).

Generators (all outward-oriented, deterministic):
  icosphere(level)        20*4^level triangles on the unit sphere
  cube_sphere(m)          6*m^2 quads (gnomonic cube projected to the sphere)
  geodesic(nu)            20*nu^2 triangles (icosahedron faces split nu-fold)
  perturbed_sphere(...)   radial perturbation by real spherical harmonics
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import capi


class Mesh:
    """Nodes and surface elements of one body (conn padded to 4 with -1)."""

    def __init__(self, nodes: np.ndarray, conn: np.ndarray, nverts: np.ndarray):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int64)
        self.nverts = np.ascontiguousarray(nverts, dtype=np.int32)

    @property
    def num_elements(self) -> int:
        return int(self.conn.shape[0])

    def collocation_points(self, corrections: int = 0) -> np.ndarray:
        """(n, 8) records x y z nx ny nz area wt, one per element
        (elements.rs:172-181, preprocess.rs:126-153), computed by libccab;
        corrections & capi.FIX_QUAD_AREA gives quads their full area."""
        out = np.zeros((self.num_elements, 8), dtype=np.float64)
        rc = capi.lib().ccab_collocation_points_ex(
            self.nodes.ctypes.data_as(ct.c_void_p), self.conn.ctypes.data_as(ct.c_void_p),
            self.nverts.ctypes.data_as(ct.c_void_p), self.num_elements,
            int(corrections) & capi.FIX_QUAD_AREA, out.ctypes.data_as(ct.c_void_p))
        if rc != capi.OK:
            raise capi.CcabError(rc, "invalid or degenerate element")
        return out


def read_vtk(path: str, body_index: int) -> Mesh:
    nodes, cells, types = [], [], []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        w = lines[i].split()
        if w and w[0] == "POINTS":
            npts = int(w[1])
            nodes = [[float(v) for v in lines[i + 1 + t].split()[:3]] for t in range(npts)]
            i += npts
        elif w and w[0] == "CELLS":
            nel = int(w[1])
            for t in range(nel):
                ww = lines[i + 1 + t].split()
                cells.append((int(ww[0]), [int(v) for v in ww[1:]]))
            i += nel
        elif w and w[0] == "CELL_TYPES":
            nel = int(w[1])
            types = [int(lines[i + 1 + t].split()[0]) for t in range(nel)]
            i += nel
        i += 1
    nbodies = 0
    sel = []
    for e, (body, ids) in enumerate(cells):
        if body > nbodies:
            if body != nbodies + 1:
                raise ValueError(f"body ids must appear as 1,2,3..; got {body} after {nbodies}")
            nbodies = body
        if body == body_index:
            sel.append((ids, types[e]))
    if body_index < 1 or body_index > nbodies:
        raise ValueError(f"body {body_index} not in mesh ({nbodies} bodies)")
    conn = np.full((len(sel), 4), -1, dtype=np.int64)
    nv = np.zeros(len(sel), dtype=np.int32)
    for e, (ids, t) in enumerate(sel):
        if t == 5:
            nv[e] = 3
        elif t == 9:
            nv[e] = 4
        else:
            raise ValueError(f"cell type {t} is not a linear surface element")
        conn[e, : nv[e]] = ids[: nv[e]]
    return Mesh(np.asarray(nodes, dtype=np.float64), conn, nv)


def write_vtk(path: str, mesh: Mesh, body_index: int = 3):
    """Write `mesh` as body `body_index` (lower bodies get one dummy vertex cell each)."""
    lines = ["# vtk DataFile Version 2.0", "cc_abstemious_amd mesh", "ASCII",
             "DATASET UNSTRUCTURED_GRID", f"POINTS {len(mesh.nodes)} double"]
    lines += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in mesh.nodes]
    cells, types = [], []
    for b in range(1, body_index):
        cells.append(f"{b} 0")
        types.append(1)
    for e in range(mesh.num_elements):
        nv = int(mesh.nverts[e])
        cells.append(f"{body_index} " + " ".join(str(int(v)) for v in mesh.conn[e, :nv]))
        types.append(5 if nv == 3 else 9)
    size = sum(len(c.split()) for c in cells)
    lines.append(f"CELLS {len(cells)} {size}")
    lines += cells
    lines.append(f"CELL_TYPES {len(types)}")
    lines += [str(t) for t in types]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def _orient_outward(nodes, faces):
    p0, p1, p2 = nodes[faces[:, 0]], nodes[faces[:, 1]], nodes[faces[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    c = (p0 + p1 + p2) / 3.0
    flip = np.einsum("ij,ij->i", nrm, c) < 0
    faces = faces.copy()
    faces[flip, 1], faces[flip, 2] = faces[flip, 2].copy(), faces[flip, 1].copy()
    return faces


def _tri_mesh(nodes, faces, radius=1.0) -> Mesh:
    faces = _orient_outward(nodes, faces)
    conn = np.full((len(faces), 4), -1, dtype=np.int64)
    conn[:, :3] = faces
    return Mesh(nodes * radius, conn, np.full(len(faces), 3, dtype=np.int32))


def icosphere(level: int, radius: float = 1.0) -> Mesh:
    """Icosahedron refined `level` times by midpoint splitting (20*4^level triangles)."""
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (np.asarray(verts[a]) + np.asarray(verts[b])) / 2.0
                m /= np.linalg.norm(m)
                cache[key] = len(verts)
                verts.append(tuple(m))
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.asarray(nf, dtype=np.int64)
    return _tri_mesh(np.asarray(verts, dtype=np.float64), f, radius)


def geodesic(nu: int, radius: float = 1.0) -> Mesh:
    """Class-I geodesic sphere: every icosahedron face split into nu^2 triangles."""
    v, f = _icosahedron()
    verts = {}
    pts = []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in verts:
            verts[key] = len(pts)
            pts.append(p / np.linalg.norm(p))
        return verts[key]

    faces = []
    for a, b, c in f:
        A, B, C = v[a], v[b], v[c]
        grid = {}
        for i in range(nu + 1):
            for j in range(nu + 1 - i):
                p = A + (B - A) * (i / nu) + (C - A) * (j / nu)
                grid[(i, j)] = vid(p)
        for i in range(nu):
            for j in range(nu - i):
                faces.append([grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]])
                if j < nu - i - 1:
                    faces.append([grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]])
    return _tri_mesh(np.asarray(pts, dtype=np.float64), np.asarray(faces, dtype=np.int64), radius)


def cube_sphere(m: int, radius: float = 1.0) -> Mesh:
    """6*m^2 quads: an m x m grid on each cube face, projected to the sphere."""
    pts = {}
    nodes = []

    def nid(p):
        key = tuple(np.round(p, 12))
        if key not in pts:
            pts[key] = len(nodes)
            nodes.append(p / np.linalg.norm(p))
        return pts[key]

    quads = []
    s = np.tan(np.linspace(-np.pi / 4, np.pi / 4, m + 1))  # equiangular grid
    for axis in range(3):
        for sign in (-1.0, 1.0):
            for a in range(m):
                for b in range(m):
                    ids = []
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = np.zeros(3)
                        p[axis] = sign
                        p[(axis + 1) % 3] = s[a + da]
                        p[(axis + 2) % 3] = s[b + db]
                        ids.append(nid(p))
                    quads.append(ids)
    nodes = np.asarray(nodes, dtype=np.float64)
    quads = np.asarray(quads, dtype=np.int64)
    # outward orientation from the first triangle of each quad
    p0, p1, p3 = nodes[quads[:, 0]], nodes[quads[:, 1]], nodes[quads[:, 3]]
    nrm = np.cross(p1 - p0, p3 - p0)
    c = nodes[quads].mean(axis=1)
    flip = np.einsum("ij,ij->i", nrm, c) < 0
    quads[flip] = quads[flip][:, ::-1]
    nv = np.full(len(quads), 4, dtype=np.int32)
    return Mesh(nodes * radius, quads, nv)


def real_sph_harm(lmax: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Orthonormal real spherical harmonics Y_lm, l <= lmax, shape ((lmax+1)^2, npts).
    theta = polar angle, phi = azimuth.  Associated Legendre by recurrence."""
    x = np.cos(theta)
    sx = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P = {}
    P[(0, 0)] = np.ones_like(x)
    for m in range(1, lmax + 1):
        P[(m, m)] = -(2 * m - 1) * sx * P[(m - 1, m - 1)]
    for m in range(0, lmax):
        P[(m + 1, m)] = (2 * m + 1) * x * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * x * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    out = []
    from math import factorial, pi, sqrt
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            K = sqrt((2 * l + 1) / (4 * pi) * factorial(l - am) / factorial(l + am))
            if m == 0:
                out.append(K * P[(l, 0)])
            elif m > 0:
                out.append(sqrt(2) * K * np.cos(m * phi) * P[(l, am)])
            else:
                out.append(sqrt(2) * K * np.sin(am * phi) * P[(l, am)])
    return np.asarray(out)


def perturbed_sphere(nu: int, coeffs: np.ndarray, lmax: int = 8) -> Mesh:
    """Geodesic sphere (20 nu^2 triangles) with r = 1 + sum c_lm Y_lm."""
    base = geodesic(nu)
    p = base.nodes
    r0 = np.linalg.norm(p, axis=1)
    theta = np.arccos(np.clip(p[:, 2] / r0, -1.0, 1.0))
    phi = np.arctan2(p[:, 1], p[:, 0])
    Y = real_sph_harm(lmax, theta, phi)
    r = 1.0 + np.asarray(coeffs, dtype=np.float64) @ Y
    nodes = p / r0[:, None] * r[:, None]
    faces = base.conn[:, :3]
    return _tri_mesh(nodes, faces, 1.0)
