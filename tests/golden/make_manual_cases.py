"""Writes manual_cases.json: the HeatFlow manual's chapter-6 known-answer cases.

Each case is a 3x1x1 steady-state system (isteady=1) with unit cell size, unit density and heat
capacity, boundary temperature 200 and boundary conductivities (x, y, z) = (1, 0, 0).  The
expected H matrix, boundary vector B, heating Q and temperatures T are the values printed in the
manual (docs/Heatflow_Manual.pdf, sections 6.1-6.4).
"""
import json
import os

CASES = [
    dict(name="case1_uniform", kappa=[1, 1, 1], qdens_centre=0.0,
         H=[[-2, 1, 0], [1, -2, 1], [0, 1, -2]], B=[200, 0, 200], Q=[0, 0, 0],
         T=[200, 200, 200]),
    dict(name="case2_uniform_heater", kappa=[1, 1, 1], qdens_centre=10.0,
         H=[[-2, 1, 0], [1, -2, 1], [0, 1, -2]], B=[200, 0, 200], Q=[0, 10, 0],
         T=[205, 210, 205]),
    dict(name="case3_hetero", kappa=[1, 2, 1], qdens_centre=0.0,
         H=[[-7 / 3, 4 / 3, 0], [4 / 3, -8 / 3, 4 / 3], [0, 4 / 3, -7 / 3]], B=[200, 0, 200],
         Q=[0, 0, 0], T=[200, 200, 200]),
    dict(name="case4_hetero_heater", kappa=[1, 2, 1], qdens_centre=10.0,
         H=[[-7 / 3, 4 / 3, 0], [4 / 3, -8 / 3, 4 / 3], [0, 4 / 3, -7 / 3]], B=[200, 0, 200],
         Q=[0, 10, 0], T=[205, 208.75, 205]),
]

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "manual_cases.json")
    with open(out, "w") as f:
        json.dump(dict(t_bath=200.0, kappa_bound=[1, 1, 0, 0, 0, 0], cases=CASES), f, indent=1)
    print(out)
