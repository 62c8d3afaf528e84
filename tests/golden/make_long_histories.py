"""Oracle runs of the benchmark configurations over their full step counts,
kept as fixtures for the GPU's full-run parity tests.

    python tests/golden/make_long_histories.py CASE [threads] [steps]

CASE is one of
  cyl_re40_200x200x3     Re 40 cylinder, 200 x 200, 3 levels, dt = 0.004,
                         impulsive start from rest, 1000 steps
  cyl_re100_800x400x5    Re 100 cylinder, 800 x 400, 5 levels, dt = 0.002,
                         Gaussian-vortex start, 5000 steps
  cyl_re200_1024x1024x5  Re 200 cylinder, 1024 x 1024, 5 levels, dt = 0.001,
                         Gaussian-vortex start, 2000 steps (of the 10000)

The oracle assembles and inverts its own B (get_Binv).  Writes
oracle_long_CASE.npz with the CD and CL of every step, the norms of Γ and ψ
per level every 500 steps (and at the end) and the final F̃b.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in ("oracle", "ibpm.jl_amd", "tests"):
    sys.path.insert(0, os.path.join(ROOT, p))

import oracle as O  # noqa: E402
from cases import boundary_of, case, gaussian_vortex_gamma, oracle_bodies, sc7  # noqa: E402

STEPS = {"cyl_re40_200x200x3": 1000, "cyl_re100_800x400x5": 5000, "cyl_re200_1024x1024x5": 2000}
VORTEX = {"cyl_re40_200x200x3": False, "cyl_re100_800x400x5": True, "cyl_re200_1024x1024x5": True}


def main(name, threads=8, nsteps=None):
    nsteps = nsteps or STEPS[name]
    O.set_threads(threads)
    c = case(name)
    g = c["grid"]
    t0 = time.time()
    P = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c))
    setup = time.time() - t0
    print(f"{name}: setup {setup:.0f}s", flush=True)
    S = P.new_state(sc7(c, 0.0))
    if VORTEX[name]:
        psi, q, G2 = P.vort2flux(gaussian_vortex_gamma(g))
        S.G[:], S.psi[:], S.q[:] = G2, psi, q
    cd = np.zeros(nsteps)
    cl = np.zeros(nsteps)
    norm_steps, nG, npsi = [], [], []
    done = 0
    while done < nsteps:
        n = min(500, nsteps - done)
        ts = (np.arange(n) + done) * c["dt"]
        a, b = P.run(S, np.stack([sc7(c, t) for t in ts]))
        cd[done:done + n], cl[done:done + n] = a[:, 0], b[:, 0]
        done += n
        norm_steps.append(done)
        nG.append(np.linalg.norm(S.G, axis=0))
        npsi.append(np.linalg.norm(S.psi, axis=0))
        print(f"{name}: {done} steps, CD {cd[done - 1]:.10f} CL {cl[done - 1]:.3e}, {time.time() - t0:.0f}s",
              flush=True)
    np.savez_compressed(os.path.join(HERE, f"oracle_long_{name}.npz"), CD=cd, CL=cl, norm_steps=np.array(norm_steps),
                        norm_G=np.array(nG), norm_psi=np.array(npsi), Fb=S.Fb.copy(), dt=c["dt"],
                        threads=threads, setup_s=setup, run_s=time.time() - t0 - setup)


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 8,
         int(sys.argv[3]) if len(sys.argv) > 3 else None)
