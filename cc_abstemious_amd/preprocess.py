"""Problem data of one analysis -- the host-side mirror of PreData
(src/preprocess.rs:24-107) and of preprocess() (preprocess.rs:110-124), which
builds it from a UserInput (src/preprocess/input_data.rs:79-91).

Collocation points live in HBM as the (n, 8) float64 record x y z nx ny nz
area wt; the frequency is implicit state selected with set_frequency_index,
as in the reference.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import capi, input_data
from .mesh import Mesh, read_vtk


def linspace(start: float, end: float, npts: int) -> list[float]:
    """tools.rs:1-10 (end point exact; one point gives [end])."""
    if npts < 1:
        raise ValueError("linspace needs at least one point")
    dx = (end - start) / (npts - 1) if npts > 1 else 0.0
    x = [start] * npts
    for i in range(1, npts - 1):
        x[i] = start + dx * i
    x[npts - 1] = end
    return x


def logspace(start: float, end: float, npts: int) -> list[float]:
    """tools.rs:12-24."""
    xl = linspace(math.log10(start), math.log10(end), npts)
    x = [start] * npts
    x[npts - 1] = end
    for i in range(1, npts - 1):
        x[i] = 10.0 ** xl[i]
    return x


@dataclasses.dataclass
class IncidentWave:
    kind: int = capi.WAVE_PLANE           # WAVE_PLANE / WAVE_SPHERICAL
    vec: tuple = (1.0, 0.0, 0.0)          # direction / origin
    amplitude: tuple = (1.0, 0.0)         # pressure amplitude

    def to_c(self) -> capi.Wave:
        return capi.wave(self.kind, self.vec, self.amplitude)


@dataclasses.dataclass
class Solver:
    """Solver of input_data.rs:27-32 plus the mixed-precision variants."""
    kind: int = capi.SOLVER_DIRECT
    tolerance: float = 1e-5
    max_iterations: int = 1000
    restart: int = 0                 # 0: ceil(sqrt(max_iterations)) as the reference
    ref_mgs: bool = True             # reference Arnoldi coefficient (gmres.rs:168)
    stop_rule: int = capi.STOP_BACKWARD_ERROR
    matrix_prec: int = capi.C128
    ir_tolerance: float = 1e-12
    ir_max_steps: int = 10
    leaf_size: int = 32              # hierarchical: solve.rs:91,117
    aca_tolerance: float = 1e-4      # hierarchical: solve.rs:92,118
    eta: float = 4.0                 # hierarchical: h_matrix.rs:136
    # ITERATIVE_MIXED with the C32 operator: FP64 residuals against the c128
    # matrix (C128) or the FP16 operator plus its complex-FP32 remainder (C32)
    residual_matrix: int = capi.C128

    def to_c(self) -> capi.SolverCfg:
        g = capi.gmres_params(self.max_iterations, self.tolerance, self.restart, self.ref_mgs,
                              self.stop_rule, self.matrix_prec)
        return capi.solver_cfg(self.kind, g, self.ir_tolerance, self.ir_max_steps,
                               self.leaf_size, self.aca_tolerance, self.eta,
                               self.residual_matrix)


class PreData:
    def __init__(self, mesh: Mesh, frequencies, sound_speed=343.0, mass_density=1.21,
                 method=capi.CLASSICAL, problem_type=capi.EXTERIOR,
                 bc_type=capi.BC_NORMAL_VELOCITY, bc_value=(0.0, 0.0),
                 incident_wave: IncidentWave | None = None, solver: Solver | None = None,
                 device: int = 0, corrections: int = 0):
        # corrections: capi.FIX_* bits (0 = the reference's numerics)
        self.corrections = int(corrections)
        self.mesh = mesh
        self.cpts = mesh.collocation_points(self.corrections)
        self.frequency_list = list(frequencies)
        self.sound_speed = float(sound_speed)
        self.mass_density = float(mass_density)
        self.method = int(method)
        self.problem_type = int(problem_type)
        self.bc_type = int(bc_type)
        self.bc_value = (float(bc_value[0]), float(bc_value[1]))
        self.incident_wave = incident_wave or IncidentWave()
        self.solver = solver or Solver()
        self.ifreq = 0
        self.device = torch.device("cuda", device)
        self.d_cpts = torch.from_numpy(self.cpts).to(self.device)
        self.field_points = np.zeros((0, 3))   # preprocess.rs get_field_points
        self.d_fps = None
        self.output_total = True               # OutputType::Total (solve.rs:176)
        self.output_pressure = True            # OutputField::Pressure (postprocess.rs:72)
        self.output_file = ""                  # Output::file (input_data.rs:73)

    def set_field_points(self, fps):
        self.field_points = np.ascontiguousarray(np.asarray(fps, dtype=np.float64).reshape(-1, 3))
        self.d_fps = torch.from_numpy(self.field_points).to(self.device)

    def get_field_points(self):
        return self.field_points

    # preprocess.rs:35-45
    def set_frequency_index(self, index: int):
        self.ifreq = int(index)

    def get_frequencies(self):
        return self.frequency_list

    def get_frequency(self) -> float:
        return self.frequency_list[self.ifreq]

    def get_angular_frequency(self) -> float:
        return 2.0 * math.pi * self.get_frequency()

    def get_wavenumber(self) -> float:
        return self.get_angular_frequency() / self.sound_speed

    def get_num_eqn(self) -> int:
        return int(self.cpts.shape[0])

    # the remaining getters of preprocess.rs:43-107
    def get_sound_speed(self) -> float:
        return self.sound_speed

    def get_mass_density(self) -> float:
        return self.mass_density

    def get_problem_type(self) -> int:
        return self.problem_type

    def get_method_type(self) -> int:
        return self.method

    def get_hdiag(self) -> complex:
        """-1/2 on the H diagonal for exterior problems (preprocess.rs:61-67)."""
        return complex(-0.5 if self.problem_type == capi.EXTERIOR else 0.0, 0.0)

    def get_gdiag(self) -> complex:
        """beta/2 = i/(2k) on the G diagonal for Burton-Miller (preprocess.rs:69-75)."""
        return 0.5j / self.get_wavenumber() if self.method == capi.BURTON_MILLER else 0j

    def get_solver(self) -> Solver:
        return self.solver

    def get_incident_wave(self) -> IncidentWave:
        return self.incident_wave

    def get_surface_bc(self):
        """(bc_type, value) of the surface boundary condition."""
        return self.bc_type, self.bc_value

    def get_mesh(self) -> Mesh:
        return self.mesh

    def get_cpts(self) -> np.ndarray:
        return self.cpts

    def get_output_filename(self) -> str:
        return self.output_file

    def get_output_field(self) -> str:
        return "Pressure" if self.output_pressure else "VelocityPotential"

    def get_output_type(self) -> str:
        return "Total" if self.output_total else "Scattered"

    def physics(self, frequency: float | None = None) -> capi.Physics:
        f = self.get_frequency() if frequency is None else frequency
        return capi.physics(f, self.sound_speed, self.mass_density, self.method,
                            self.problem_type, self.bc_type, self.bc_value, self.corrections)

    def physics_array(self, frequencies=None):
        fs = self.frequency_list if frequencies is None else list(frequencies)
        arr = (capi.Physics * len(fs))()
        for i, f in enumerate(fs):
            arr[i] = self.physics(f)
        return arr


_METHOD = {"Classical": capi.CLASSICAL, "BurtonMiller": capi.BURTON_MILLER}
_PROBLEM = {"Exterior": capi.EXTERIOR, "Interior": capi.INTERIOR}
_BC = {"Pressure": capi.BC_PRESSURE, "NormalVelocity": capi.BC_NORMAL_VELOCITY,
       "Impedance": capi.BC_IMPEDANCE}
_WAVE = {"PlaneWave": capi.WAVE_PLANE, "SphericalWave": capi.WAVE_SPHERICAL}
_SOLVER = {"Direct": capi.SOLVER_DIRECT, "Iterative": capi.SOLVER_ITERATIVE,
           "Hierarchical": capi.SOLVER_HIERARCHICAL}


def solver_from_input(s: input_data.SolverInput) -> Solver:
    """Solver enum of input_data.rs:27-32 as a solver configuration: Direct is
    the FP64 LU, Iterative and Hierarchical keep the reference's GMRES (restart
    ceil(sqrt(max_iterations)), backward-error stop, Arnoldi coefficient of
    gmres.rs:168) and, for Hierarchical, leaf 32 and ACA tolerance 1e-4
    (solve.rs:91-92)."""
    if s.kind == "Direct":
        return Solver(kind=capi.SOLVER_DIRECT)
    return Solver(kind=_SOLVER[s.kind], tolerance=s.tolerance, max_iterations=s.max_iterations)


def preprocess(user_input: input_data.UserInput, device: int = 0) -> PreData:
    """preprocess.rs:110-124: read the mesh body, expand the frequency list and
    compute the collocation points of the analysis described by user_input."""
    u = user_input
    m = read_vtk(u.mesh_file, u.body_index)
    pd = PreData(m, input_data.frequency_list(u.frequency), u.sound_speed, u.mass_density,
                 _METHOD[u.method_type], _PROBLEM[u.problem_type], _BC[u.surface_bc.bc_type],
                 u.surface_bc.value,
                 IncidentWave(_WAVE[u.incident_wave.kind], tuple(u.incident_wave.vec),
                              tuple(u.incident_wave.amplitude)),
                 solver_from_input(u.solver), device=device)
    pd.output_total = u.output.o_type == "Total"
    pd.output_pressure = u.output.field == "Pressure"
    pd.output_file = u.output.file
    pd.set_field_points(np.asarray(u.output.field_points, dtype=np.float64).reshape(-1, 3))
    pd.input = u
    return pd


def colmajor(n: int, m: int, dtype, device) -> torch.Tensor:
    """An n x m column-major tensor (strides (1, n)): data_ptr is the BLAS base."""
    return torch.empty((m, n), dtype=dtype, device=device).t()
