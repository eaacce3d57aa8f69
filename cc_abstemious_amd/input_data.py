"""JSON input of an analysis -- mirror of src/preprocess/input_data.rs.

The reference deserialises ``UserInput`` with serde's default, externally
tagged enum representation (input_data.rs:8-91):

* unit variants are strings: ``"problem_type": "Exterior"``;
* struct variants are one-key objects: ``"solver": {"Iterative":
  {"tolerance": 1e-5, "max_iterations": 1000}}``, ``"solver": {"Direct": {}}``;
* fixed arrays have exactly their length (``[f64; 3]``, ``[f64; 2]``);
* integer fields (``usize``) take JSON integers only; float fields take any
  JSON number;
* a missing field or an unknown variant is an error; unknown keys are ignored.

``read_input_json`` / ``read_input_string`` (input_data.rs:93-108) raise
``InputError`` where the reference's ``Result`` carries a serde error.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os


class InputError(ValueError):
    """Malformed analysis input (the Err of read_input_json)."""


# --- primitive field readers ----------------------------------------------

def _field(obj: dict, name: str, where: str):
    if not isinstance(obj, dict):
        raise InputError(f"{where}: expected an object")
    if name not in obj:
        raise InputError(f"{where}: missing field `{name}`")
    return obj[name]


def _f64(v, where: str) -> float:
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise InputError(f"{where}: expected f64, got {v!r}")
    return float(v)


def _usize(v, where: str) -> int:
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise InputError(f"{where}: expected usize, got {v!r}")
    return int(v)


def _string(v, where: str) -> str:
    if not isinstance(v, str):
        raise InputError(f"{where}: expected a string, got {v!r}")
    return v


def _f64_array(v, n: int, where: str) -> tuple:
    if not isinstance(v, list) or len(v) != n:
        raise InputError(f"{where}: expected an array of {n} f64")
    return tuple(_f64(x, where) for x in v)


def _unit_variant(v, variants: tuple, where: str) -> str:
    if isinstance(v, str) and v in variants:
        return v
    raise InputError(f"{where}: unknown variant {v!r}, expected one of {', '.join(variants)}")


def _struct_variant(v, variants: tuple, where: str):
    """(tag, body) of an externally tagged struct variant."""
    if not isinstance(v, dict) or len(v) != 1:
        raise InputError(f"{where}: expected an object with exactly one variant key")
    tag, body = next(iter(v.items()))
    if tag not in variants:
        raise InputError(f"{where}: unknown variant {tag!r}, expected one of {', '.join(variants)}")
    if not isinstance(body, dict):
        raise InputError(f"{where}.{tag}: expected an object")
    return tag, body


# --- the input types (input_data.rs:8-91) -----------------------------------

@dataclasses.dataclass
class FrequencyInput:
    """List {values} | LinearSpaced {start, end, number} | LogSpaced {...}
    (input_data.rs:8-13)."""
    kind: str = "List"
    values: list = dataclasses.field(default_factory=list)
    start: float = 0.0
    end: float = 0.0
    number: int = 0

    VARIANTS = ("List", "LinearSpaced", "LogSpaced")

    @classmethod
    def from_json(cls, v):
        tag, b = _struct_variant(v, cls.VARIANTS, "frequency")
        if tag == "List":
            vals = _field(b, "values", "frequency.List")
            if not isinstance(vals, list):
                raise InputError("frequency.List.values: expected an array")
            return cls(kind=tag, values=[_f64(x, "frequency.List.values") for x in vals])
        w = f"frequency.{tag}"
        return cls(kind=tag, start=_f64(_field(b, "start", w), w + ".start"),
                   end=_f64(_field(b, "end", w), w + ".end"),
                   number=_usize(_field(b, "number", w), w + ".number"))

    def to_json(self):
        if self.kind == "List":
            return {"List": {"values": list(self.values)}}
        return {self.kind: {"start": self.start, "end": self.end, "number": self.number}}


@dataclasses.dataclass
class SolverInput:
    """Direct {} | Iterative {tolerance, max_iterations} |
    Hierarchical {tolerance, max_iterations} (input_data.rs:27-32)."""
    kind: str = "Direct"
    tolerance: float = 0.0
    max_iterations: int = 0

    VARIANTS = ("Direct", "Iterative", "Hierarchical")

    @classmethod
    def from_json(cls, v):
        tag, b = _struct_variant(v, cls.VARIANTS, "solver")
        if tag == "Direct":
            return cls(kind=tag)
        w = f"solver.{tag}"
        return cls(kind=tag, tolerance=_f64(_field(b, "tolerance", w), w + ".tolerance"),
                   max_iterations=_usize(_field(b, "max_iterations", w), w + ".max_iterations"))

    def to_json(self):
        if self.kind == "Direct":
            return {"Direct": {}}
        return {self.kind: {"tolerance": self.tolerance, "max_iterations": self.max_iterations}}


@dataclasses.dataclass
class IncidentWaveInput:
    """PlaneWave {direction, amplitude} | SphericalWave {origin, amplitude}
    (input_data.rs:34-44)."""
    kind: str = "PlaneWave"
    vec: tuple = (1.0, 0.0, 0.0)          # direction or origin
    amplitude: tuple = (0.0, 0.0)

    VARIANTS = ("PlaneWave", "SphericalWave")

    @classmethod
    def from_json(cls, v):
        tag, b = _struct_variant(v, cls.VARIANTS, "incident_wave")
        w = f"incident_wave.{tag}"
        key = "direction" if tag == "PlaneWave" else "origin"
        return cls(kind=tag, vec=_f64_array(_field(b, key, w), 3, f"{w}.{key}"),
                   amplitude=_f64_array(_field(b, "amplitude", w), 2, w + ".amplitude"))

    def to_json(self):
        key = "direction" if self.kind == "PlaneWave" else "origin"
        return {self.kind: {key: list(self.vec), "amplitude": list(self.amplitude)}}


@dataclasses.dataclass
class SurfaceBoundaryCondition:
    """bc_type: Pressure | NormalVelocity | Impedance; value [re, im]
    (input_data.rs:46-57)."""
    bc_type: str = "NormalVelocity"
    value: tuple = (0.0, 0.0)

    VARIANTS = ("Pressure", "NormalVelocity", "Impedance")

    @classmethod
    def from_json(cls, v):
        return cls(bc_type=_unit_variant(_field(v, "bc_type", "surface_bc"), cls.VARIANTS,
                                         "surface_bc.bc_type"),
                   value=_f64_array(_field(v, "value", "surface_bc"), 2, "surface_bc.value"))

    def to_json(self):
        return {"bc_type": self.bc_type, "value": list(self.value)}


@dataclasses.dataclass
class Output:
    """o_type: Total | Scattered; field: Pressure | VelocityPotential; file;
    field_points (input_data.rs:59-77)."""
    o_type: str = "Scattered"
    field: str = "Pressure"
    file: str = ""
    field_points: list = dataclasses.field(default_factory=list)

    @classmethod
    def from_json(cls, v):
        fps = _field(v, "field_points", "output")
        if not isinstance(fps, list):
            raise InputError("output.field_points: expected an array")
        return cls(o_type=_unit_variant(_field(v, "o_type", "output"), ("Total", "Scattered"),
                                        "output.o_type"),
                   field=_unit_variant(_field(v, "field", "output"),
                                       ("Pressure", "VelocityPotential"), "output.field"),
                   file=_string(_field(v, "file", "output"), "output.file"),
                   field_points=[_f64_array(p, 3, "output.field_points") for p in fps])

    def to_json(self):
        return {"o_type": self.o_type, "field": self.field, "file": self.file,
                "field_points": [list(p) for p in self.field_points]}


@dataclasses.dataclass
class UserInput:
    """input_data.rs:79-91."""
    mesh_file: str = ""
    body_index: int = 3
    frequency: FrequencyInput = dataclasses.field(default_factory=FrequencyInput)
    sound_speed: float = 1.0
    mass_density: float = 1.0
    problem_type: str = "Exterior"
    method_type: str = "Classical"
    solver: SolverInput = dataclasses.field(default_factory=SolverInput)
    incident_wave: IncidentWaveInput = dataclasses.field(default_factory=IncidentWaveInput)
    surface_bc: SurfaceBoundaryCondition = dataclasses.field(
        default_factory=SurfaceBoundaryCondition)
    output: Output = dataclasses.field(default_factory=Output)

    @classmethod
    def from_json(cls, v):
        w = "UserInput"
        return cls(
            mesh_file=_string(_field(v, "mesh_file", w), "mesh_file"),
            body_index=_usize(_field(v, "body_index", w), "body_index"),
            frequency=FrequencyInput.from_json(_field(v, "frequency", w)),
            sound_speed=_f64(_field(v, "sound_speed", w), "sound_speed"),
            mass_density=_f64(_field(v, "mass_density", w), "mass_density"),
            problem_type=_unit_variant(_field(v, "problem_type", w), ("Interior", "Exterior"),
                                       "problem_type"),
            method_type=_unit_variant(_field(v, "method_type", w), ("Classical", "BurtonMiller"),
                                      "method_type"),
            solver=SolverInput.from_json(_field(v, "solver", w)),
            incident_wave=IncidentWaveInput.from_json(_field(v, "incident_wave", w)),
            surface_bc=SurfaceBoundaryCondition.from_json(_field(v, "surface_bc", w)),
            output=Output.from_json(_field(v, "output", w)))

    def to_json(self) -> dict:
        return {"mesh_file": self.mesh_file, "body_index": self.body_index,
                "frequency": self.frequency.to_json(), "sound_speed": self.sound_speed,
                "mass_density": self.mass_density, "problem_type": self.problem_type,
                "method_type": self.method_type, "solver": self.solver.to_json(),
                "incident_wave": self.incident_wave.to_json(),
                "surface_bc": self.surface_bc.to_json(), "output": self.output.to_json()}


def _reject_nonfinite_literals(token: str):
    # serde_json has no NaN / Infinity literals
    raise InputError(f"invalid JSON number {token}")


def read_input_string(text: str) -> UserInput:
    """input_data.rs:105-108."""
    try:
        obj = json.loads(text, parse_constant=_reject_nonfinite_literals)
    except json.JSONDecodeError as e:
        raise InputError(f"invalid JSON: {e}") from None
    return UserInput.from_json(obj)


def read_input_json(path) -> UserInput:
    """input_data.rs:93-103 (the path is taken relative to the working directory)."""
    with open(os.fspath(path), encoding="utf-8") as f:
        return read_input_string(f.read())


def frequency_list(freq: FrequencyInput) -> list:
    """process_frequency_list (preprocess.rs:155-168)."""
    from .preprocess import linspace, logspace
    if freq.kind == "List":
        return list(freq.values)
    if freq.number < 1:
        raise InputError(f"frequency.{freq.kind}.number must be at least 1")
    if freq.kind == "LinearSpaced":
        return linspace(freq.start, freq.end, freq.number)
    if freq.start <= 0.0 or freq.end <= 0.0 or not math.isfinite(freq.start * freq.end):
        raise InputError("frequency.LogSpaced: start and end must be positive")
    return logspace(freq.start, freq.end, freq.number)
