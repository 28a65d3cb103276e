"""Writes reference_known_answers.json: the inputs and pass thresholds of
the reference's own tests that pin this path, transcribed as data with the
file:line they come from (paths relative to the OzW1234/huji-rich tree).
The reference stores no golden output vectors for the path; these are its
known-answer inputs and grader tolerances."""
import json
import os

KNOWN = {
    "hydrostatic_convection_mass_flux": {
        "source": "tests/newtonian/common/hydrostatic_convection_mass_flux/test.cpp:13-21, test.py:7",
        "density": 1.2, "pressure": 3.4, "energy": 5.6, "sound_speed": 7.8,
        "face_velocity": 9.1, "expected_mass_flux": "-density*face_velocity (equality)",
    },
    "hll_acoustic_eulerian": {
        "source": "tests/newtonian/common/hll_acoustic_eulerian/test.cpp:44-56, test.py:58-96",
        "density": 2.0, "pressure": 3.0, "dp": 1e-2, "gamma": 5.0 / 3.0, "threshold": 0.01,
    },
    "hll_advection_eulerian": {
        "source": "tests/newtonian/common/hll_advection_eulerian/test.cpp:44-56, test.py:58-83",
        "left_density": 1.0, "right_density": 2.0, "pressure": 3.0, "velocity": 4.0,
        "gamma": 5.0 / 3.0, "threshold": 0.01,
    },
    "gresho_vortex_2o": {
        "source": "tests/newtonian/two_dimensional/gresho_vortex_2o/test.cpp:140-231, test.py:115-118",
        "box": [-0.5, 0.5, -0.5, 0.5], "n": 30, "gamma": 5.0 / 3.0, "cfl": 0.3, "t_end": 0.003,
        "l1_density": 2e-4, "l1_pressure": 2e-3, "l1_radial_velocity": 5e-4,
        "l1_azimuthal_velocity": 1e-3,
    },
    "conservation_eulerian": {
        "source": "tests/newtonian/two_dimensional/conservation_eulerian/test.cpp:82-191, test.py:3-34",
        "note": "first order, rigid walls, static mesh: totals of mass, momentum and energy "
                "equal every step (the grader compares 6 significant digits)",
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_known_answers.json")
    with open(out, "w") as f:
        json.dump(KNOWN, f, indent=1, sort_keys=True)
    print("wrote", out)
