"""Field results and their CSV files -- mirror of src/postprocess.rs.

FPResult                                   postprocess.rs:6-11
write_results_at_frequency(predata, r)     postprocess.rs:13-43
write_results_at_point(predata, rs, i)     postprocess.rs:45-68
convert_results(predata, rs)               postprocess.rs:70-81 (+ 83-88)

The CSV files are byte-compatible with the reference's: the csv crate's
default writer (comma, records ended by a single '\\n', no quoting of these
fields) with every number formatted as Rust's ``Display`` prints it
(``rust_display``).
"""
from __future__ import annotations

import dataclasses
import decimal
import math

import numpy as np


@dataclasses.dataclass
class FPResult:
    """One frequency's field-point values (host arrays): ``scattered`` holds the
    total field when the output type is Total (postprocess.rs:6-11)."""
    frequency: float
    scattered: np.ndarray | None = None
    incident: np.ndarray | None = None
    radiated_power: float = 0.0


def rust_display(x: float) -> str:
    """A float as Rust's ``f64::to_string`` writes it: the shortest digits
    that read back to the same double, in positional notation (never an
    exponent), without a trailing ``.0``; ``NaN``, ``inf``, ``-inf``, ``-0``."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    s = format(decimal.Decimal(repr(x)), "f")
    if "." in s:
        s = s.rstrip("0").rstrip(".")
    return s


def _write_csv(path: str, rows):
    with open(path, "w", newline="") as f:
        for row in rows:
            f.write(",".join(row))
            f.write("\n")


def _field_name(predata) -> str:
    return "pre" if predata.get_output_field() == "Pressure" else "vpo"


def write_results_at_frequency(predata, result: FPResult):
    """All field points at one frequency: index, x, y, z, <field>_re,
    <field>_im, inc_re, inc_im (postprocess.rs:13-43)."""
    scatt = np.asarray(result.scattered)
    inc = np.asarray(result.incident)
    points = predata.get_field_points()
    if scatt.shape[0] != len(points):
        raise ValueError("Error while writing results: pressure array and field point array "
                         "not equal length!")
    on = _field_name(predata)
    rows = [["index", "x", "y", "z", f"{on}_re", f"{on}_im", "inc_re", "inc_im"]]
    for i in range(scatt.shape[0]):
        rows.append([str(i), rust_display(points[i][0]), rust_display(points[i][1]),
                     rust_display(points[i][2]), rust_display(scatt[i].real),
                     rust_display(scatt[i].imag), rust_display(inc[i].real),
                     rust_display(inc[i].imag)])
    _write_csv(predata.get_output_filename(), rows)


def write_results_at_point(predata, results, index: int):
    """One field point over all frequencies: freq, <field>_re, <field>_im,
    inc_re, inc_im (postprocess.rs:45-68)."""
    on = _field_name(predata)
    rows = [["freq", f"{on}_re", f"{on}_im", "inc_re", "inc_im"]]
    for r in results:
        s, c = complex(r.scattered[index]), complex(r.incident[index])
        rows.append([rust_display(r.frequency), rust_display(s.real), rust_display(s.imag),
                     rust_display(c.real), rust_display(c.imag)])
    _write_csv(predata.get_output_filename(), rows)


def convert_results(predata, results):
    """Velocity potential to pressure, p = i omega rho phi, on the scattered
    and incident arrays of every result when the output field is Pressure
    (postprocess.rs:70-88).  For results built from potentials by the caller:
    the field kernel behind Analysis.run applies the same factor itself."""
    if predata.get_output_field() != "Pressure":
        return
    rho = predata.get_mass_density()
    for r in results:
        scale = complex(0.0, 2.0 * math.pi * r.frequency * rho)
        r.scattered = np.asarray(r.scattered) * scale
        r.incident = np.asarray(r.incident) * scale
