"""Write naca64.json: the random parameters of the batched NACA 0012 workload.

64 independent simulations of a unit-chord NACA 0012 on (-1, 3) x (-1, 1):
finest grids 400x200 (22 instances), 600x300 (21) and 800x400 (21); levels
uniform in {3, 4, 5}; angle of attack uniform in [0, 20] degrees, applied as
the freestream inclination (timestepping.jl:259,267-268); Re log-uniform in
[100, 1000]; dt = 0.25 dx.  Drawn once from numpy.random.default_rng(0) and
committed, so no run has to reproduce the generator.

    python tests/golden/make_naca64.py
"""
import json
import os

import numpy as np


def main():
    rng = np.random.default_rng(0)
    sizes = [(400, 200)] * 22 + [(600, 300)] * 21 + [(800, 400)] * 21
    out = []
    for k, (nx, ny) in enumerate(sizes):
        mg = int(rng.integers(3, 6))
        alpha_deg = float(rng.uniform(0.0, 20.0))
        Re = float(10 ** rng.uniform(2.0, 3.0))
        dx = 4.0 / nx
        out.append({"nx": nx, "ny": ny, "mg": mg, "alpha_deg": alpha_deg, "Re": Re, "dx": dx, "dt": 0.25 * dx})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "naca64.json")
    with open(path, "w") as f:
        json.dump({"box": [-1.0, 3.0, -1.0, 1.0], "airfoil": "0012", "chord": 1.0, "instances": out}, f, indent=1)
    print(path, len(out))


if __name__ == "__main__":
    main()
