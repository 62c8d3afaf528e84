"""Run the CPU oracle on the reference's steady flat-plate benchmark and record
its drag coefficient next to the Julia number it pins against.

    python tests/golden/run_oracle_pins.py [threads]

Case: benchmarks/eldredge/steady_bl.jl with benchmarks/eldredge/config.jl --
400 x 200 cells on (-1.5, 6.5) x (-2, 2), mg = 5, Re = 100, plate of unit chord
at zero incidence through the origin with 51 points, dt = 0.01, advance! at
t = 0:0.01:100 (10001 steps).  Julia's final CD: 0.39415052
(benchmarks/eldredge/README.md:25).  Writes oracle_steady_bl.json.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "ibpm.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from cases import boundary_of, case, oracle_bodies, sc7  # noqa: E402


def main(threads=8, nsteps=10001):
    O.set_threads(threads)
    c = case("steady_bl")
    g = c["grid"]
    t0 = time.time()
    P = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c))
    setup = time.time() - t0
    S = P.new_state(sc7(c, 0.0))
    sc = sc7(c, 0.0)
    hist = []
    t0 = time.time()
    chunk = 100
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        cd, cl = P.run(S, np.tile(sc, (n, 1)))
        done += n
        hist.append([done, float(cd[-1, 0]), float(cl[-1, 0])])
        print(done, cd[-1, 0], cl[-1, 0], f"{time.time() - t0:.0f}s", flush=True)
    out = {"case": "steady_bl", "steps": nsteps, "CD_final": float(S.CD[0]), "CL_final": float(S.CL[0]),
           "CD_julia": 0.39415052, "rel_diff": abs(float(S.CD[0]) - 0.39415052) / 0.39415052,
           "history_every_100": hist, "threads": threads, "setup_s": setup, "run_s": time.time() - t0}
    with open(os.path.join(HERE, "oracle_steady_bl.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "history_every_100"}))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8, int(sys.argv[2]) if len(sys.argv) > 2 else 10001)
