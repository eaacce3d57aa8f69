"""Host-side checks of the analysis input, output and driver plumbing (no GPU):
the reference's JSON input (input_data.rs:110-130 and its input_1.json), the
frequency lists (tools.rs:1-25), Rust's float formatting used by the CSV
writers (postprocess.rs:13-68) and the closed form the GPU analysis tests
compare with (oracle/analytic.py)."""
import json
import math
import os
import random

import numpy as np
import pytest

from cc_abstemious_amd import input_data, postprocess
from cc_abstemious_amd.input_data import InputError, read_input_json, read_input_string
from cc_abstemious_amd.preprocess import linspace, logspace
from oracle import analytic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_read_input_json_fixture():
    """input_data.rs:113-119 (json_reader): body_index of src/tests/input_1.json."""
    u = read_input_json(os.path.join(GOLD, "input_1.json"))
    assert u.body_index == 3
    assert u.mesh_file == "./src/tests/sphere_refined_2.vtk"
    assert u.frequency.kind == "List" and u.frequency.values == [100.0]
    assert (u.sound_speed, u.mass_density) == (1500.0, 1000.0)
    assert (u.problem_type, u.method_type) == ("Exterior", "Classical")
    assert u.solver.kind == "Direct"
    assert u.incident_wave.kind == "PlaneWave"
    assert u.incident_wave.vec == (0.0, 1.0, 0.0) and u.incident_wave.amplitude == (1.0, 0.0)
    assert u.surface_bc.bc_type == "NormalVelocity" and u.surface_bc.value == (0.0, 0.0)
    assert (u.output.o_type, u.output.field) == ("Scattered", "Pressure")
    assert u.output.field_points == [(0.0, 1.0, 1.0)]


def _base():
    with open(os.path.join(GOLD, "input_1.json")) as f:
        return json.load(f)


def test_input_round_trip_and_variants():
    obj = _base()
    obj["solver"] = {"Hierarchical": {"tolerance": 1e-5, "max_iterations": 1000}}
    obj["frequency"] = {"LogSpaced": {"start": 10, "end": 1000.0, "number": 7}}
    obj["incident_wave"] = {"SphericalWave": {"origin": [0.1, 0, 0], "amplitude": [2, -1]}}
    obj["surface_bc"] = {"bc_type": "Impedance", "value": [415.03, 83.006]}
    obj["output"]["o_type"] = "Total"
    obj["output"]["field"] = "VelocityPotential"
    obj["method_type"] = "BurtonMiller"
    obj["problem_type"] = "Interior"
    obj["unknown_key"] = {"ignored": True}          # serde ignores unknown fields
    u = read_input_string(json.dumps(obj))
    assert u.solver.kind == "Hierarchical" and u.solver.max_iterations == 1000
    assert u.frequency.kind == "LogSpaced" and u.frequency.start == 10.0
    assert u.incident_wave.kind == "SphericalWave" and u.incident_wave.amplitude == (2.0, -1.0)
    u2 = read_input_string(json.dumps(u.to_json()))
    assert u2 == u
    obj["solver"] = {"Direct": {}}
    assert read_input_string(json.dumps(obj)).solver.kind == "Direct"


@pytest.mark.parametrize("mutate,msg", [
    (lambda o: o.pop("mesh_file"), "missing field `mesh_file`"),
    (lambda o: o.__setitem__("body_index", 3.0), "expected usize"),
    (lambda o: o.__setitem__("body_index", -1), "expected usize"),
    (lambda o: o.__setitem__("sound_speed", "1500"), "expected f64"),
    (lambda o: o.__setitem__("problem_type", "Outside"), "unknown variant"),
    (lambda o: o.__setitem__("solver", {"Gauss": {}}), "unknown variant"),
    (lambda o: o.__setitem__("solver", {"Iterative": {"tolerance": 1e-5}}),
     "missing field `max_iterations`"),
    (lambda o: o.__setitem__("solver", {"Direct": {}, "Iterative": {}}), "exactly one"),
    (lambda o: o["incident_wave"]["PlaneWave"].__setitem__("direction", [1, 0]), "array of 3"),
    (lambda o: o["output"]["field_points"].append([1, 2]), "array of 3"),
    (lambda o: o["surface_bc"].__setitem__("bc_type", "Rigid"), "unknown variant"),
    (lambda o: o["frequency"].__setitem__("List", {"values": 100.0}), "expected an array"),
])
def test_input_errors(mutate, msg):
    obj = _base()
    mutate(obj)
    with pytest.raises(InputError, match=msg):
        read_input_string(json.dumps(obj))


def test_input_rejects_bad_json():
    with pytest.raises(InputError, match="invalid JSON"):
        read_input_string("{")
    text = json.dumps(_base()).replace('"sound_speed": 1500.0', '"sound_speed": NaN')
    with pytest.raises(InputError):
        read_input_string(text)


def test_frequency_lists():
    """tools.rs:1-25: end points exact, interior points start + dx*i and
    10^(linspace of the logs)."""
    x = linspace(10.0, 1000.0, 50)
    assert x[0] == 10.0 and x[-1] == 1000.0
    dx = 990.0 / 49
    assert all(x[i] == 10.0 + dx * i for i in range(1, 49))
    np.testing.assert_allclose(x, np.linspace(10.0, 1000.0, 50), rtol=1e-15)
    y = logspace(10.0, 1000.0, 5)
    assert y[0] == 10.0 and y[-1] == 1000.0
    np.testing.assert_allclose(y, [10.0, 10 ** 1.5, 100.0, 10 ** 2.5, 1000.0], rtol=1e-14)
    assert linspace(3.0, 7.0, 1) == [7.0]          # one point: x[npts-1] = end
    f = input_data.FrequencyInput(kind="LinearSpaced", start=27.3, end=436.7, number=64)
    assert input_data.frequency_list(f) == linspace(27.3, 436.7, 64)
    with pytest.raises(InputError):
        input_data.frequency_list(input_data.FrequencyInput(kind="LinearSpaced", number=0))


@pytest.mark.parametrize("x,s", [
    (1.0, "1"), (100.0, "100"), (-2.5, "-2.5"), (0.1, "0.1"), (-0.0, "-0"), (0.0, "0"),
    (1e-7, "0.0000001"), (1.5e-5, "0.000015"), (1e21, "1000000000000000000000"),
    (123456789.125, "123456789.125"), (0.30000000000000004, "0.30000000000000004"),
    (-3.9380091293979903e-05, "-0.0000393800912939799"),   # 15 digits read it back
    (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"),
])
def test_rust_display(x, s):
    """Rust's Display for f64: shortest round-trip digits, positional."""
    assert postprocess.rust_display(x) == s


def test_rust_display_round_trips():
    rng = random.Random(5)
    for _ in range(2000):
        x = rng.uniform(-1, 1) * 10.0 ** rng.randint(-30, 30)
        s = postprocess.rust_display(x)
        assert "e" not in s and float(s) == x
    assert postprocess.rust_display(5e-324) == "0." + "0" * 323 + "5"


class _StubPreData:
    def __init__(self, path, field="Pressure", fps=None):
        self.path, self.field = path, field
        self.fps = fps if fps is not None else [(10.0, 0.0, 0.0), (0.0, -2.5, 1.0)]

    def get_field_points(self):
        return self.fps

    def get_output_field(self):
        return self.field

    def get_output_filename(self):
        return self.path

    def get_mass_density(self):
        return 1000.0


def test_csv_writers(tmp_path):
    """postprocess.rs:13-68: header, one record per point (or frequency),
    comma separated, records ended by '\\n'."""
    path = str(tmp_path / "out.csv")
    pd = _StubPreData(path)
    r0 = postprocess.FPResult(100.0, np.array([-3.9380091293979903e-05 + 4.7605686656319906e-04j,
                                               1.0 - 2.0j]),
                              np.array([0.5 + 0j, -0.25 + 1e-7j]))
    r1 = postprocess.FPResult(150.5, np.array([2.0 + 0j, 3j]), np.array([1j, 1.0 + 0j]))
    postprocess.write_results_at_frequency(pd, r0)
    with open(path, "rb") as f:
        data = f.read()
    assert data == (b"index,x,y,z,pre_re,pre_im,inc_re,inc_im\n"
                    b"0,10,0,0,-0.0000393800912939799,0.00047605686656319906,0.5,0\n"
                    b"1,0,-2.5,1,1,-2,-0.25,0.0000001\n")
    pd.field = "VelocityPotential"
    postprocess.write_results_at_point(pd, [r0, r1], 1)
    with open(path, "rb") as f:
        data = f.read()
    assert data == (b"freq,vpo_re,vpo_im,inc_re,inc_im\n"
                    b"100,1,-2,-0.25,0.0000001\n"
                    b"150.5,0,3,1,0\n")
    with pytest.raises(ValueError, match="not equal length"):
        postprocess.write_results_at_frequency(_StubPreData(path, fps=[(1.0, 2.0, 3.0)]), r0)


def test_convert_results():
    """postprocess.rs:70-88: p = i omega rho phi on both arrays, Pressure only."""
    r = postprocess.FPResult(50.0, np.array([1.0 + 2.0j]), np.array([-1.0j]))
    postprocess.convert_results(_StubPreData("x", field="VelocityPotential"), [r])
    assert r.scattered[0] == 1.0 + 2.0j
    postprocess.convert_results(_StubPreData("x"), [r])
    w = 2 * math.pi * 50.0 * 1000.0
    assert r.scattered[0] == (1.0 + 2.0j) * complex(0.0, w)
    assert r.incident[0] == -1.0j * complex(0.0, w)


@pytest.mark.parametrize("ka", [0.2, 1.0, 4.0])
def test_rigid_sphere_series(ka):
    """The closed form satisfies the rigid-surface condition
    d/dr (p_inc - p_s) = 0 at r = a (the reference's sign of p_s), and has
    converged in its number of terms."""
    a = 0.5
    k = ka / a
    for ct in (1.0, 0.3, -0.7):
        f = lambda r: np.exp(1j * k * r * ct) - analytic.rigid_sphere_scattered(a, k, r, ct)
        h = 1e-6
        d = (f(a + h) - f(a - h)) / (2 * h)
        assert abs(d) < 1e-8 * k
        # a sign error in p_s leaves d of the size of the incident derivative
        g = lambda r: np.exp(1j * k * r * ct) + analytic.rigid_sphere_scattered(a, k, r, ct)
        assert abs((g(a + h) - g(a - h)) / (2 * h)) > 1e-3 * k
    p40 = analytic.rigid_sphere_scattered(a, k, 3.0, 0.3, nterms=40)
    p80 = analytic.rigid_sphere_scattered(a, k, 3.0, 0.3, nterms=80)
    assert abs(p40 - p80) <= 1e-14 * abs(p80)


def test_cli_usage():
    from cc_abstemious_amd.__main__ import main
    assert main([]) == 2
