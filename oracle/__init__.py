"""CPU FP64 oracle for the cc_abstemious collocation-BEM hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import it;
the product (``cc_abstemious_amd``) never does.  It wraps ``liboracle.so``
(built from ``ccab_oracle.c`` by ``make -C oracle``), a C restatement of the
reference Rust crate; ``ccab_oracle.h`` lists the reference file:line each
function follows and which golden vector pins it.

Also holds an independent ASCII-VTK reader that follows the reference's
``Mesh::read_from_vtk`` (src/preprocess/mesh_data.rs:96-175) including its
body-id quirk (the first integer of each CELLS line is the body id).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

CLASSICAL, BURTON_MILLER = 0, 1
EXTERIOR, INTERIOR = 0, 1
BC_PRESSURE, BC_NORMAL_VELOCITY, BC_IMPEDANCE = 0, 1, 2
SOLVER_DIRECT, SOLVER_ITERATIVE, SOLVER_HIERARCHICAL = 0, 1, 2
WAVE_PLANE, WAVE_SPHERICAL = 0, 1


class Physics(ct.Structure):
    _fields_ = [
        ("frequency", ct.c_double),
        ("sound_speed", ct.c_double),
        ("mass_density", ct.c_double),
        ("method", ct.c_int32),
        ("problem_type", ct.c_int32),
        ("bc_type", ct.c_int32),
        ("pad_", ct.c_int32),
        ("bc_value", ct.c_double * 2),
    ]


class Wave(ct.Structure):
    _fields_ = [
        ("kind", ct.c_int32),
        ("pad_", ct.c_int32),
        ("vec", ct.c_double * 3),
        ("amplitude", ct.c_double * 2),
    ]


class Blocks(ct.Structure):
    _fields_ = [
        ("nblocks", ct.c_int64),
        ("row_off", ct.POINTER(ct.c_int64)),
        ("col_off", ct.POINTER(ct.c_int64)),
        ("rows", ct.POINTER(ct.c_int64)),
        ("cols", ct.POINTER(ct.c_int64)),
        ("admissible", ct.POINTER(ct.c_int32)),
        ("root_diameter", ct.c_double),
    ]


class Aca(ct.Structure):
    _fields_ = [
        ("m", ct.c_int64),
        ("n", ct.c_int64),
        ("rank", ct.c_int64),
        ("u", ct.c_void_p),
        ("v", ct.c_void_p),
        ("norm", ct.c_double),
        ("rank_before_recompression", ct.c_int64),
    ]


ROWCOL_FN = ct.CFUNCTYPE(None, ct.c_void_p, ct.c_int, ct.c_int64, ct.c_void_p)
RC_FN = ct.CFUNCTYPE(None, ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64,
                     ct.c_void_p)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        P = ct.c_void_p
        i64 = ct.c_int64
        L.orc_collocation_points.argtypes = [P, P, P, i64, P]
        L.orc_gh_entry.argtypes = [ct.POINTER(Physics), P, i64, i64, P, P]
        L.orc_dense_surface_matrices.argtypes = [ct.POINTER(Physics), P, i64, P, P, ct.c_int,
                                                 ct.c_int]
        L.orc_dense_field_matrices.argtypes = [ct.POINTER(Physics), P, i64, P, i64, P, P]
        L.orc_surface_row_or_column.argtypes = [ct.POINTER(Physics), P, ct.c_int, P, i64, P,
                                                i64, P]
        L.orc_incident_surface.argtypes = [ct.POINTER(Physics), ct.POINTER(Wave), P, i64, P]
        L.orc_incident_field.argtypes = [ct.POINTER(Physics), ct.POINTER(Wave), P, i64, P]
        L.orc_lu_solve.argtypes = [i64, P, P]
        L.orc_lu_factor.argtypes = [i64, P, i64, P]
        L.orc_lu_apply.argtypes = [i64, P, i64, P, P]
        L.orc_gmres.argtypes = [i64, P, P, i64, ct.c_double, ct.c_int, P, P]
        L.orc_solve_surface_dense.argtypes = [ct.POINTER(Physics), P, i64, P, ct.c_int, i64,
                                              ct.c_double, P, P, P]
        L.orc_field.argtypes = [ct.POINTER(Physics), P, i64, P, i64, P, P, P, ct.c_int,
                                ct.c_int, P]
        L.orc_generate_random_ab.argtypes = [i64, i64, ct.c_uint64, P, P]
        L.orc_build_blocks.argtypes = [P, i64, i64, i64, ct.c_double, ct.POINTER(Blocks)]
        L.orc_blocks_free.argtypes = [ct.POINTER(Blocks)]
        L.orc_block_tree_admissible_at.argtypes = [P, i64, i64, i64, ct.c_double, P, ct.c_int]
        L.orc_aca_build.argtypes = [ct.c_double, i64, i64, ROWCOL_FN, P, ct.POINTER(Aca)]
        L.orc_aca_free.argtypes = [ct.POINTER(Aca)]
        L.orc_hmatrix_build.argtypes = [i64, RC_FN, P, P, i64, i64, ct.c_double, ct.c_double]
        L.orc_hmatrix_build.restype = P
        L.orc_hmatrix_free.argtypes = [P]
        L.orc_hmatrix_norm.argtypes = [P]
        L.orc_hmatrix_norm.restype = ct.c_double
        L.orc_hmatrix_stats.argtypes = [P, P]
        L.orc_hmatrix_to_full.argtypes = [P, P]
        L.orc_solve_surface_hmatrix.argtypes = [ct.POINTER(Physics), P, i64, P, i64,
                                                ct.c_double, i64, ct.c_double, ct.c_double,
                                                P, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def physics(frequency, sound_speed=343.0, mass_density=1.21, method=CLASSICAL,
            problem_type=EXTERIOR, bc_type=BC_NORMAL_VELOCITY, bc_value=(0.0, 0.0)) -> Physics:
    ph = Physics()
    ph.frequency = frequency
    ph.sound_speed = sound_speed
    ph.mass_density = mass_density
    ph.method = method
    ph.problem_type = problem_type
    ph.bc_type = bc_type
    ph.bc_value[0], ph.bc_value[1] = bc_value
    return ph


def wave(kind=WAVE_PLANE, vec=(1.0, 0.0, 0.0), amplitude=(1.0, 0.0)) -> Wave:
    w = Wave()
    w.kind = kind
    for d in range(3):
        w.vec[d] = vec[d]
    w.amplitude[0], w.amplitude[1] = amplitude
    return w


# --------------------------------------------------------------------------
# Mesh reading (mesh_data.rs:96-175) and geometry (elements.rs, preprocess.rs)
# --------------------------------------------------------------------------
def read_vtk(path: str):
    """Return (nodes[nn,3], elements list of (body_id, node_ids, vtk_type))."""
    nodes = []
    cells = []
    types = []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        words = lines[i].split()
        if words and words[0] == "POINTS":
            npts = int(words[1])
            for t in range(npts):
                nodes.append([float(v) for v in lines[i + 1 + t].split()[:3]])
            i += npts
        elif words and words[0] == "CELLS":
            nel = int(words[1])
            for t in range(nel):
                w = lines[i + 1 + t].split()
                cells.append((int(w[0]), [int(v) for v in w[1:]]))
            i += nel
        elif words and words[0] == "CELL_TYPES":
            nel = int(words[1])
            for t in range(nel):
                types.append(int(lines[i + 1 + t].split()[0]))
            i += nel
        i += 1
    elems = [(b, ids, types[e]) for e, (b, ids) in enumerate(cells)]
    return np.asarray(nodes, dtype=np.float64), elems


def body_elements(elems, body_index: int):
    """Elements of body `body_index` (1-based) in file order; bodies must
    appear with ids 1,2,3... (mesh_data.rs:163-171)."""
    nb = 0
    out = []
    for b, ids, t in elems:
        if b > nb:
            nb += 1
        if b == body_index:
            out.append((ids, t))
    return out


def collocation_points(nodes: np.ndarray, conn: np.ndarray, nverts: np.ndarray) -> np.ndarray:
    """(nelem, 8) array x y z nx ny nz area wt (one centroid point per element)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    nverts = np.ascontiguousarray(nverts, dtype=np.int32)
    ne = conn.shape[0]
    out = np.zeros((ne, 8), dtype=np.float64)
    rc = lib().orc_collocation_points(_p(nodes), _p(conn), _p(nverts), ne, _p(out))
    if rc != 0:
        raise ValueError("unsupported element type")
    return out


def mesh_cpts_from_vtk(path: str, body_index: int) -> np.ndarray:
    nodes, elems = read_vtk(path)
    sel = body_elements(elems, body_index)
    conn = np.full((len(sel), 4), -1, dtype=np.int64)
    nv = np.zeros(len(sel), dtype=np.int32)
    for e, (ids, t) in enumerate(sel):
        if t == 5:
            nv[e] = 3
        elif t == 9:
            nv[e] = 4
        else:
            raise ValueError(f"element type {t} is not a surface element")
        conn[e, : nv[e]] = ids[: nv[e]]
    return collocation_points(nodes, conn, nv)


# --------------------------------------------------------------------------
# Thin wrappers
# --------------------------------------------------------------------------
def dense_surface_matrices(ph: Physics, cpts: np.ndarray, nthreads: int = 0, locked: bool = False):
    cpts = np.ascontiguousarray(cpts, dtype=np.float64)
    n = cpts.shape[0]
    H = np.zeros((n, n), dtype=np.complex128, order="F")
    G = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().orc_dense_surface_matrices(ct.byref(ph), _p(cpts), n, _p(H), _p(G), nthreads,
                                     int(locked))
    return H, G


def gh_entry(ph: Physics, cpts, i, j):
    g = np.zeros(1, np.complex128)
    h = np.zeros(1, np.complex128)
    lib().orc_gh_entry(ct.byref(ph), _p(np.ascontiguousarray(cpts)), i, j, _p(g), _p(h))
    return g[0], h[0]


def surface_row_or_column(ph: Physics, cpts, side: int, i, j) -> np.ndarray:
    ii = np.ascontiguousarray(np.atleast_1d(i), dtype=np.int64)
    jj = np.ascontiguousarray(np.atleast_1d(j), dtype=np.int64)
    out = np.zeros(max(len(ii), len(jj)), dtype=np.complex128)
    rc = lib().orc_surface_row_or_column(ct.byref(ph), _p(np.ascontiguousarray(cpts)), side,
                                         _p(ii), len(ii), _p(jj), len(jj), _p(out))
    if rc != 0:
        raise ValueError("i or j must be a singleton")
    return out


def incident_surface(ph: Physics, w: Wave, cpts) -> np.ndarray:
    cpts = np.ascontiguousarray(cpts, dtype=np.float64)
    out = np.zeros(cpts.shape[0], dtype=np.complex128)
    lib().orc_incident_surface(ct.byref(ph), ct.byref(w), _p(cpts), cpts.shape[0], _p(out))
    return out


def incident_field(ph: Physics, w: Wave, fps) -> np.ndarray:
    fps = np.ascontiguousarray(fps, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(fps.shape[0], dtype=np.complex128)
    lib().orc_incident_field(ct.byref(ph), ct.byref(w), _p(fps), fps.shape[0], _p(out))
    return out


def lu_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.complex128)
    x = _c128(b).copy()
    lib().orc_lu_solve(A.shape[0], _p(A), _p(x))
    return x


def gmres(A: np.ndarray, b: np.ndarray, max_it: int, thresh: float, ref_mgs: bool = True):
    A = np.asfortranarray(A, dtype=np.complex128)
    x = _c128(b).copy()
    stats = np.zeros(3, dtype=np.int64)
    err = np.zeros(1, dtype=np.float64)
    lib().orc_gmres(A.shape[0], _p(A), _p(x), max_it, thresh, int(ref_mgs), _p(stats), _p(err))
    return x, {"num_mv": int(stats[0]), "iterations": int(stats[1]), "restarts": int(stats[2]),
               "backward_error": float(err[0])}


def solve_surface_dense(ph: Physics, cpts, rhs_inc, solver=SOLVER_DIRECT, max_it=1000,
                        tol=1e-5):
    cpts = np.ascontiguousarray(cpts, dtype=np.float64)
    n = cpts.shape[0]
    phi = np.zeros(n, np.complex128)
    vn = np.zeros(n, np.complex128)
    stats = np.zeros(3, dtype=np.int64)
    lib().orc_solve_surface_dense(ct.byref(ph), _p(cpts), n, _p(_c128(rhs_inc)), solver, max_it,
                                  tol, _p(phi), _p(vn), _p(stats))
    return phi, vn, stats


def solve_surface_hmatrix(ph: Physics, cpts, rhs_inc, max_it=1000, tol=1e-5, leaf=32,
                          aca_tol=1e-4, eta=4.0):
    cpts = np.ascontiguousarray(cpts, dtype=np.float64)
    n = cpts.shape[0]
    phi = np.zeros(n, np.complex128)
    vn = np.zeros(n, np.complex128)
    stats = np.zeros(7, dtype=np.int64)
    lib().orc_solve_surface_hmatrix(ct.byref(ph), _p(cpts), n, _p(_c128(rhs_inc)), max_it, tol,
                                    leaf, aca_tol, eta, _p(phi), _p(vn), _p(stats))
    return phi, vn, stats


def field(ph: Physics, cpts, fps, phi, vn, phi_inc_fp, want_total=False, to_pressure=True):
    cpts = np.ascontiguousarray(cpts, dtype=np.float64)
    fps = np.ascontiguousarray(fps, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(fps.shape[0], np.complex128)
    lib().orc_field(ct.byref(ph), _p(cpts), cpts.shape[0], _p(fps), fps.shape[0],
                    _p(_c128(phi)), _p(_c128(vn)), _p(_c128(phi_inc_fp)), int(want_total),
                    int(to_pressure), _p(out))
    return out


def generate_random_ab(m: int, n: int, seed: int):
    a = np.zeros((m, n), dtype=np.complex128, order="F")
    b = np.zeros(n, dtype=np.complex128)
    lib().orc_generate_random_ab(m, n, seed, _p(a), _p(b))
    return a, b


def build_blocks(coords: np.ndarray, leaf: int, eta: float):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    B = Blocks()
    lib().orc_build_blocks(_p(coords), coords.shape[1], coords.shape[0], leaf, eta, ct.byref(B))
    out = []
    for b in range(B.nblocks):
        r = [B.rows[t] for t in range(B.row_off[b], B.row_off[b + 1])]
        c = [B.cols[t] for t in range(B.col_off[b], B.col_off[b + 1])]
        out.append((r, c, bool(B.admissible[b])))
    diam = B.root_diameter
    lib().orc_blocks_free(ct.byref(B))
    return out, diam


def block_tree_admissible_at(coords, leaf, eta, path):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    p = np.ascontiguousarray(path, dtype=np.int32)
    return lib().orc_block_tree_admissible_at(_p(coords), coords.shape[1], coords.shape[0], leaf,
                                              eta, _p(p), len(path))


def aca(tol: float, A: np.ndarray):
    """ACA of an explicit matrix (the reference's mult_aca test harness)."""
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape

    def cb(user, is_row, idx, out):
        src = np.ascontiguousarray(A[idx, :] if is_row else A[:, idx])
        ct.memmove(out, src.ctypes.data, src.nbytes)

    fn = ROWCOL_FN(cb)
    res = Aca()
    lib().orc_aca_build(tol, m, n, fn, None, ct.byref(res))
    r = res.rank
    U = np.ctypeslib.as_array(ct.cast(res.u, ct.POINTER(ct.c_double)), shape=(2 * m * r,)).copy()
    V = np.ctypeslib.as_array(ct.cast(res.v, ct.POINTER(ct.c_double)), shape=(2 * n * r,)).copy()
    U = U.view(np.complex128).reshape(r, m).T
    V = V.view(np.complex128).reshape(r, n).T
    out = {"U": U, "V": V, "norm": res.norm, "rank": r,
           "rank_before": res.rank_before_recompression}
    lib().orc_aca_free(ct.byref(res))
    return out


class HMatrix:
    """Oracle H-matrix over an explicit matrix or a callback (h_matrix.rs)."""

    def __init__(self, n, rowcol, coords, leaf, tol, eta=4.0):
        coords = np.ascontiguousarray(coords, dtype=np.float64)

        def cb(user, rows, nr, cols, nc, out):
            r = np.ctypeslib.as_array(ct.cast(rows, ct.POINTER(ct.c_int64)), shape=(nr,))
            c = np.ctypeslib.as_array(ct.cast(cols, ct.POINTER(ct.c_int64)), shape=(nc,))
            vals = np.ascontiguousarray(rowcol(r.copy(), c.copy()), dtype=np.complex128)
            ct.memmove(out, vals.ctypes.data, vals.nbytes)

        self._fn = RC_FN(cb)
        self.n = n
        self._h = lib().orc_hmatrix_build(n, self._fn, None, _p(coords), coords.shape[1], leaf,
                                          tol, eta)

    def gemv(self, alpha, x, beta, b):
        b = _c128(b).copy()
        a = np.array([alpha], np.complex128)
        bb = np.array([beta], np.complex128)
        # pass complex scalars by value through a struct
        lib().orc_hmatrix_gemv.argtypes = [ct.c_void_p, _C128, ct.c_void_p, _C128, ct.c_void_p]
        lib().orc_hmatrix_gemv(self._h, _C128(a.real[0], a.imag[0]), _p(_c128(x)),
                               _C128(bb.real[0], bb.imag[0]), _p(b))
        return b

    def norm(self):
        return lib().orc_hmatrix_norm(self._h)

    def stats(self):
        s = np.zeros(4, np.int64)
        lib().orc_hmatrix_stats(self._h, _p(s))
        return {"admissible": int(s[0]), "inadmissible": int(s[1]), "storage": int(s[2]),
                "rank_sum": int(s[3]),
                "compression_ratio": 100.0 * (1.0 - s[2] / float(self.n * self.n))}

    def to_full(self):
        A = np.zeros((self.n, self.n), np.complex128, order="F")
        lib().orc_hmatrix_to_full(self._h, _p(A))
        return A

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_hmatrix_free(self._h)
            self._h = None


class _C128(ct.Structure):
    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]
