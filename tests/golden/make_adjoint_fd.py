"""Central finite differences of the CPU oracle for the pitching-plate
gradient, kept as the fixture the GPU adjoint is checked against.

    python tests/golden/make_adjoint_fd.py [threads] [nsteps]

Case: flat plate (chord 1, 101 points, mid-chord at the origin) in the
body-fixed MovingGrid frame, Re = 200, 512 x 256 cells on [-1,3] x [-1,1],
4 levels, dt = 0.002, impulsive start from rest.  Motion: U(t) = cos(beta),
V(t) = sin(beta) (the inflow direction in the body frame), theta(t) =
A sin(2 pi f t), theta_dot = 2 pi f A cos(2 pi f t); freestream zero.
Parameters: A = 15 deg and f = 0.2, each scaled by 1 + 0.1 u with u drawn
uniformly in [-1, 1] by numpy.random.default_rng(2026); beta = 0.
Objective: J = mean of CD (state.CD after each step) over the last half of
the steps (steps nsteps/2 .. nsteps-1, 0-based).
Writes adjoint_fd_plate.json: the parameters, J, and dJ/d(A, f, beta) by
central differences with the steps given there.
"""
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in ("oracle", "ibpm.jl_amd", "tests"):
    sys.path.insert(0, os.path.join(ROOT, p))

import oracle as O  # noqa: E402
from cases import boundary_of, case, oracle_bodies  # noqa: E402


def params():
    u = np.random.default_rng(2026).uniform(-1.0, 1.0, 2)
    return 15.0 * math.pi / 180.0 * (1.0 + 0.1 * u[0]), 0.2 * (1.0 + 0.1 * u[1]), 0.0


def scalars(ts, A, f, beta):
    """Per-step oracle scalars {Ux, Uy, incl, U, V, theta, theta_dot}."""
    w = 2.0 * math.pi * f
    return np.array([[0.0, 0.0, 0.0, math.cos(beta), math.sin(beta), A * math.sin(w * t), w * A * math.cos(w * t)]
                     for t in ts])


def objective(P, c, nsteps, A, f, beta):
    S = P.new_state(scalars([0.0], A, f, beta)[0])
    ts = np.arange(nsteps) * c["dt"]
    cd, cl = P.run(S, scalars(ts, A, f, beta))
    return float(np.mean(cd[nsteps // 2:, 0]))


def main(threads=8, nsteps=4000):
    O.set_threads(threads)
    c = case("plate_pitch_512x256x4")
    g = c["grid"]
    t0 = time.time()
    P = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), motion=1)
    A, f, beta = params()
    J0 = objective(P, c, nsteps, A, f, beta)
    print(f"J = {J0!r} ({time.time() - t0:.0f}s)", flush=True)
    steps = {"A": 1e-4, "f": 1e-5, "beta": 1e-4}
    grad = {}
    for name, h in steps.items():
        vals = []
        for sgn in (1.0, -1.0):
            p = {"A": A, "f": f, "beta": beta}
            p[name] += sgn * h
            vals.append(objective(P, c, nsteps, p["A"], p["f"], p["beta"]))
        grad[name] = (vals[0] - vals[1]) / (2.0 * h)
        print(f"dJ/d{name} = {grad[name]!r} ({time.time() - t0:.0f}s)", flush=True)
    out = {"case": "plate_pitch_512x256x4", "nsteps": nsteps, "window": [nsteps // 2, nsteps],
           "A": A, "f": f, "beta": beta, "J": J0, "grad": grad, "fd_steps": steps,
           "threads": threads, "seconds": time.time() - t0}
    name = "adjoint_fd_plate.json" if nsteps == 4000 else f"adjoint_fd_plate_{nsteps}.json"
    with open(os.path.join(HERE, name), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8, int(sys.argv[2]) if len(sys.argv) > 2 else 4000)
