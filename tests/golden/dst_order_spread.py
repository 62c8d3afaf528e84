"""How far apart FP64 CPU codes drift when they compute the 2D DST-I
differently.  The oracle with its own transform (variant 0) is compared with
  variant 1: the same RODFT00 computed one line per FFT, dimension 2 first;
  variant 2: the half-length algorithm of the HIP library (a P = M/2 complex
             FFT and a recurrence over k for the odd outputs);
  variant 3: variant 2 with the recurrence summed with compensation.
All runs share one Binv and one initial state, so the spread comes from the
transform's rounding alone.

    python tests/golden/dst_order_spread.py [threads] [steps]

Case: Re = 200 cylinder, 1024 x 1024 cells (dx = 1/128 on [-2,6]x[-4,4]),
5 levels, dt = 1e-3, 402 body points, Gaussian-vortex start.  Writes
dst_order_spread.json: for each variant, norm-relative differences from
variant 0 of F̃b, Γ, ψ, q and the force-scaled CD/CL differences after each of
the given steps, plus F̃b after a (1/4, 1/2, 1/4) filter along the body
points.
"""
import json
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


def rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def smooth(F):
    """(1/4, 1/2, 1/4) along the closed body, per component."""
    n = F.size // 2
    out = []
    for c in (F[:n], F[n:]):
        out.append(0.25 * np.roll(c, 1) + 0.5 * c + 0.25 * np.roll(c, -1))
    return np.concatenate(out)


def main(threads=8, checkpoints=(10, 40)):
    O.set_threads(threads)
    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    t0 = time.time()
    O.set_dst_variant(0)
    P0 = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c))
    binv = P0.binv()
    print(f"Binv assembled in {time.time() - t0:.0f}s", flush=True)
    variants = (0, 1, 2, 3)
    probs = [P0] + [O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=binv)
                    for _ in variants[1:]]
    G = gaussian_vortex_gamma(g)
    states = []
    for v, P in zip(variants, probs):
        O.set_dst_variant(v)
        S = P.new_state(sc7(c, 0.0))
        psi, q, G2 = P.vort2flux(G)
        S.G[:], S.psi[:], S.q[:] = G2, psi, q
        states.append(S)
    out = {"case": "cyl_re200_1024x1024x5", "threads": threads, "checkpoints": []}
    done = 0
    for stop in checkpoints:
        n = stop - done
        ts = (np.arange(n) + done) * c["dt"]
        sc = np.stack([sc7(c, t) for t in ts])
        for v, P, S in zip(variants, probs, states):
            O.set_dst_variant(v)
            P.run(S, sc)
        done = stop
        b = states[0]
        scale = max(np.max(np.abs(b.CD)), np.max(np.abs(b.CL)))
        for v, a in zip(variants[1:], states[1:]):
            row = {"variant": v, "step": stop, "Fb": rel(a.Fb, b.Fb), "Fb_smoothed": rel(smooth(a.Fb), smooth(b.Fb)),
                   "G": rel(a.G, b.G), "psi": rel(a.psi, b.psi), "q": rel(a.q, b.q),
                   "CD": float(np.max(np.abs(a.CD - b.CD)) / scale), "CL": float(np.max(np.abs(a.CL - b.CL)) / scale)}
            out["checkpoints"].append(row)
            print(json.dumps(row), f"{time.time() - t0:.0f}s", flush=True)
    O.set_dst_variant(0)
    with open(os.path.join(HERE, "dst_order_spread.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    th = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    steps = tuple(int(x) for x in sys.argv[2].split(",")) if len(sys.argv) > 2 else (10, 40)
    main(th, steps)
