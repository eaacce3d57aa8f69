"""Problem data of one analysis -- the host-side mirror of PreData
(src/preprocess.rs:254-332) and of the UserInput fields the hot path reads
(src/preprocess/input_data.rs:471-484).

Collocation points live in HBM as the (n, 8) float64 record x y z nx ny nz
area wt; the frequency is implicit state selected with set_frequency_index,
as in the reference.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import capi
from .mesh import Mesh


def linspace(start: float, end: float, npts: int) -> list[float]:
    """tools.rs:1-10 (end point exact)."""
    dx = (end - start) / (npts - 1)
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
    """Solver of input_data.rs:419-424 plus the mixed-precision variants."""
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

    def to_c(self) -> capi.SolverCfg:
        g = capi.gmres_params(self.max_iterations, self.tolerance, self.restart, self.ref_mgs,
                              self.stop_rule, self.matrix_prec)
        return capi.solver_cfg(self.kind, g, self.ir_tolerance, self.ir_max_steps,
                               self.leaf_size, self.aca_tolerance, self.eta)


class PreData:
    def __init__(self, mesh: Mesh, frequencies, sound_speed=343.0, mass_density=1.21,
                 method=capi.CLASSICAL, problem_type=capi.EXTERIOR,
                 bc_type=capi.BC_NORMAL_VELOCITY, bc_value=(0.0, 0.0),
                 incident_wave: IncidentWave | None = None, solver: Solver | None = None,
                 device: int = 0):
        self.mesh = mesh
        self.cpts = mesh.collocation_points()
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

    def set_field_points(self, fps):
        self.field_points = np.ascontiguousarray(np.asarray(fps, dtype=np.float64).reshape(-1, 3))
        self.d_fps = torch.from_numpy(self.field_points).to(self.device)

    def get_field_points(self):
        return self.field_points

    # preprocess.rs:263-304
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

    def physics(self, frequency: float | None = None) -> capi.Physics:
        f = self.get_frequency() if frequency is None else frequency
        return capi.physics(f, self.sound_speed, self.mass_density, self.method,
                            self.problem_type, self.bc_type, self.bc_value)

    def physics_array(self, frequencies=None):
        fs = self.frequency_list if frequencies is None else list(frequencies)
        arr = (capi.Physics * len(fs))()
        for i, f in enumerate(fs):
            arr[i] = self.physics(f)
        return arr


def colmajor(n: int, m: int, dtype, device) -> torch.Tensor:
    """An n x m column-major tensor (strides (1, n)): data_ptr is the BLAS base."""
    return torch.empty((m, n), dtype=dtype, device=device).t()
