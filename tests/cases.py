"""Problem set-ups shared by the tests, the smoke check and the benchmark.

Each builder returns the reference-style objects (grid, bodies, Re, dt,
freestream) from the host mirror, plus the raw arrays the oracle takes.
"""
from __future__ import annotations

import math

import numpy as np

import ibpm


def sc7(prob_like, t):
    """Closures at t as the oracle's 7 scalars {Ux, Uy, incl, U, V, θ, θ̇}."""
    fs = prob_like["freestream"]
    s = [fs.get("Ux", lambda t: 0.0)(t), fs.get("Uy", lambda t: 0.0)(t), fs.get("inclination", lambda t: 0.0)(t),
         0.0, 0.0, 0.0, 0.0]
    m = prob_like.get("motion")
    if m is not None:
        s[3:] = [m.U(t), m.V(t), m.θ(t), m.θ̇(t)]
    return np.array(s, dtype=np.float64)


def oracle_steps(prob_like, ts):
    """Per-step scalars of the oracle: sc7, then for MotionFunction bodies
    their kinematics {x, y, θ, ẋ, ẏ, θ̇} at t."""
    rows = []
    for t in ts:
        r = [sc7(prob_like, t)]
        for b in prob_like["bodies"]:
            if isinstance(b.motion, ibpm.MotionFunction):
                r.append(np.array(b.motion.kinematics(t)))
        rows.append(np.concatenate(r))
    return np.array(rows)


def case(name: str):
    """Named configurations.  Sizes are small unless the name says otherwise."""
    U1 = {"Ux": lambda t: 1.0}
    if name == "tiny":  # 2 levels, 48 x 32, cylinder
        grid = ibpm.make_grid(0.125, (-2.0, 4.0, -2.0, 2.0), mg=2)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.1, n=24)]
        return dict(grid=grid, bodies=bodies, Re=40.0, dt=0.02, freestream=U1)
    if name == "small3":  # 3 levels, 64 x 64 (M = 64: pure radix-8 path)
        grid = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -2.0, 2.0), mg=3)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=40)]
        return dict(grid=grid, bodies=bodies, Re=100.0, dt=0.01, freestream=U1)
    if name == "mixed":  # radix 3/5 paths: nx = 120 (P = 60), ny = 40 (P = 20), 2 bodies, inclined stream
        grid = ibpm.make_grid(0.05, (-2.0, 4.0, -1.0, 1.0), mg=2)
        bodies = [ibpm.make_cylinder(0.25, grid.h, 0.0, 0.0, n=30), ibpm.make_plate(0.5, 0.3, grid.h, 1.0, 0.1, n=12)]
        return dict(grid=grid, bodies=bodies, Re=200.0, dt=0.005,
                    freestream={"Ux": lambda t: 1.0, "Uy": lambda t: 0.1, "inclination": lambda t: 0.05})
    if name == "spec":  # 256 x 200, 2 levels: compile-time transform lengths P = 128 (x) and 100 (y)
        grid = ibpm.make_grid(1.0 / 32, (-4.0, 4.0, -3.125, 3.125), mg=2)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.3, -0.2, n=100)]
        return dict(grid=grid, bodies=bodies, Re=150.0, dt=0.005, freestream=U1)
    if name == "spec2":  # 600 x 300, 2 levels: compile-time transform lengths P = 300 (x, radices 4 5 5 3) and 150 (y, 2 5 5 3)
        grid = ibpm.make_grid(0.02, (-4.0, 8.0, -3.0, 3.0), mg=2)
        bodies = [ibpm.make_cylinder(0.25, grid.h, 0.1, -0.05, n=80)]
        return dict(grid=grid, bodies=bodies, Re=120.0, dt=0.005, freestream=U1)
    if name == "direct":  # 56 x 44, 2 levels: P = 28 and 22 have factors 7 and 11, so both directions take the direct O(n^2) DST
        grid = ibpm.make_grid(0.125, (-2.0, 5.0, -2.75, 2.75), mg=2)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=26)]
        return dict(grid=grid, bodies=bodies, Re=50.0, dt=0.02, freestream=U1)
    if name == "single":  # 64 x 48, one level: no boundary terms anywhere (P = 32 and 24 = 8 x 3)
        grid = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -1.5, 1.5), mg=1)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=40)]
        return dict(grid=grid, bodies=bodies, Re=100.0, dt=0.01, freestream=U1)
    if name == "moving":  # MovingGrid pitching plate about its mid-chord at the origin
        grid = ibpm.make_grid(1.0 / 32, (-1.0, 3.0, -1.0, 1.0), mg=3)
        A, f = 15 * math.pi / 180, 0.2
        mot = ibpm.MovingGrid(lambda t: 1.0, lambda t: 0.0, lambda t: A * math.sin(2 * math.pi * f * t),
                              lambda t: 2 * math.pi * f * A * math.cos(2 * math.pi * f * t))
        bodies = [ibpm.make_plate(1.0, 0.0, grid.h, -0.5, 0.0, motion=mot, n=33)]
        return dict(grid=grid, bodies=bodies, Re=200.0, dt=0.004, freestream={}, motion=mot)
    if name == "lab":  # MotionFunction: a cylinder translating and spinning through a 3-level grid
        grid = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -2.0, 2.0), mg=3)
        w = 2 * math.pi
        mot = ibpm.MotionFunction([lambda t: 0.1 * math.sin(w * t), lambda t: 0.05 * (1 - math.cos(w * t)),
                                   lambda t: 0.4 * t],
                                  [lambda t: 0.1 * w * math.cos(w * t), lambda t: 0.05 * w * math.sin(w * t),
                                   lambda t: 0.4])
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, motion=mot, n=40)]
        return dict(grid=grid, bodies=bodies, Re=100.0, dt=0.01, freestream=U1)
    if name == "plate_pitch_512x256x4":  # pitching plate in the body-fixed frame: 512 x 256, 4 levels, nominal A = 15 deg, f = 0.2
        grid = ibpm.make_grid(1.0 / 128, (-1.0, 3.0, -1.0, 1.0), mg=4)
        A, f = 15 * math.pi / 180, 0.2
        mot = ibpm.MovingGrid(lambda t: 1.0, lambda t: 0.0, lambda t: A * math.sin(2 * math.pi * f * t),
                              lambda t: 2 * math.pi * f * A * math.cos(2 * math.pi * f * t))
        bodies = [ibpm.make_plate(1.0, 0.0, grid.h, -0.5, 0.0, motion=mot, n=101)]
        return dict(grid=grid, bodies=bodies, Re=200.0, dt=0.002, freestream={}, motion=mot)
    if name == "eldredge_lab":  # pitch_maneuver_motion.jl: plate pitching about its leading edge in the lab frame
        grid = ibpm.make_grid(0.02, (-1.5, 6.5, -2.0, 2.0), mg=5)
        a, t1, t2, t3, t4, amax = 11.0, 1.0, 3.0, 4.0, 6.0, 45 * math.pi / 180
        G = lambda t: math.log(math.cosh(a * (t - t1)) * math.cosh(a * (t - t4))  # noqa: E731
                               / (math.cosh(a * (t - t2)) * math.cosh(a * (t - t3))))
        Gd = lambda t: a * (math.tanh(a * (t - t1)) - math.tanh(a * (t - t2))  # noqa: E731
                            - math.tanh(a * (t - t3)) + math.tanh(a * (t - t4)))
        Gmax = max(G(k * 1e-3) for k in range(10001))
        mot = ibpm.MotionFunction([lambda t: 0.0, lambda t: 0.0, lambda t: -amax * G(t) / Gmax],
                                  [lambda t: 0.0, lambda t: 0.0, lambda t: -amax * Gd(t) / Gmax])
        bodies = [ibpm.make_plate(1.0, 0.0, grid.h, 0.0, 0.0, motion=mot, n=51)]
        return dict(grid=grid, bodies=bodies, Re=100.0, dt=1e-3, freestream=U1)
    if name == "cyl_re40_200x200x3":  # Re 40 cylinder, impulsive start: 200 x 200, 3 levels
        grid = ibpm.make_grid(0.02, (-1.0, 3.0, -2.0, 2.0), mg=3)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=157)]
        return dict(grid=grid, bodies=bodies, Re=40.0, dt=0.004, freestream=U1)
    if name == "steady_bl":  # benchmarks/eldredge/steady_bl.jl + config.jl
        grid = ibpm.make_grid(0.02, (-1.5, 6.5, -2.0, 2.0), mg=5)
        bodies = [ibpm.make_plate(1.0, 0.0, grid.h, 0.0, 0.0, n=51)]
        return dict(grid=grid, bodies=bodies, Re=100.0, dt=1e-2,
                    freestream={"Ux": lambda t: 1.0, "Uy": lambda t: 0.0, "inclination": lambda t: 0.0})
    if name == "cyl40_single":  # benchmarks/cyl_40.jl
        grid = ibpm.make_grid(0.02, (-1.0, 3.0, -2.0, 2.0), mg=1)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0)]
        return dict(grid=grid, bodies=bodies, Re=40.0, dt=1e-2,
                    freestream={"Ux": lambda t: 1.0, "Uy": lambda t: 0.0, "inclination": lambda t: 0.0})
    if name == "cyl_re100_800x400x5":  # Re 100 cylinder: 800 x 400, 5 levels
        grid = ibpm.make_grid(0.01, (-2.0, 6.0, -2.0, 2.0), mg=5)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=314)]
        return dict(grid=grid, bodies=bodies, Re=100.0, dt=0.002, freestream=U1)
    if name == "cyl_re200_1024x1024x5":  # Re 200 cylinder: 1024 x 1024, 5 levels, dx = 1/128
        grid = ibpm.make_grid(1.0 / 128, (-2.0, 6.0, -4.0, 4.0), mg=5)
        n = int(round(math.pi / grid.h))  # points spaced ~ dx on the unit-diameter circle
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=n)]
        return dict(grid=grid, bodies=bodies, Re=200.0, dt=0.001, freestream=U1)
    raise KeyError(name)


def naca_arclength(dx, spec="0012", chord=1.0, x_le=0.0):
    """NACA surface points equispaced in arc length (spacing close to dx),
    from the thickness law of naca() (sample-bodies.jl:104-134), leading edge
    at (x_le, 0), clockwise from the trailing edge over the upper surface.
    Deterministic: no random numbers."""
    n_fine = 20001
    xs = 0.5 * (1 - np.cos(np.linspace(0.0, np.pi, n_fine)))  # fine stations, dense at both edges
    _, xU, xL, yU, yL = ibpm.naca(xs, spec)
    # closed contour TE -> upper -> LE -> lower -> TE
    X = np.concatenate([xU[::-1], xL[1:]]) * chord + x_le
    Y = np.concatenate([yU[::-1], yL[1:]]) * chord
    seg = np.hypot(np.diff(X), np.diff(Y))
    s = np.concatenate([[0.0], np.cumsum(seg)])
    n = int(round(s[-1] / dx))
    sk = (np.arange(n) + 0.5) * (s[-1] / n)  # midpoints of n equal arcs
    xb = np.stack([np.interp(sk, s, X), np.interp(sk, s, Y)], axis=1)
    ds = np.full(n, s[-1] / n)
    return xb, ds


def naca64_cases(path=None):
    """The 64 instances of the batched NACA 0012 workload (tests/golden/naca64.json)."""
    import json
    import os
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "naca64.json")
    spec = json.load(open(path))
    x0, x1, y0, y1 = spec["box"]
    out = []
    for inst in spec["instances"]:
        grid = ibpm.make_grid(inst["dx"], (x0, x1, y0, y1), mg=inst["mg"])
        assert (grid.nx, grid.ny) == (inst["nx"], inst["ny"])
        xb, ds = naca_arclength(inst["dx"], spec["airfoil"], spec["chord"])
        body = ibpm.RigidBody(ibpm.Static(), xb, xb.copy(), np.zeros_like(xb), ds)
        a = math.radians(inst["alpha_deg"])
        out.append(dict(grid=grid, bodies=[body], Re=inst["Re"], dt=inst["dt"],
                        freestream={"Ux": lambda t: 1.0, "Uy": lambda t: 0.0, "inclination": (lambda a: lambda t: a)(a)}))
    return out


def oracle_bodies(c):
    return [(b.xb, b.ds) for b in c["bodies"]]


def boundary_of(grid):
    return (-grid.offx, -grid.offx + grid.len, -grid.offy, -grid.offy + grid.ny * grid.h)


def gaussian_vortex_gamma(grid, x0=1.5, y0=0.3, circ=0.01, width=0.1):
    """Γ_l = hc^2 ω(x, y) on every level with ω a Gaussian vortex of the given
    circulation (the seeded perturbation of the benchmark configurations)."""
    nx, ny, mg, h = grid.nx, grid.ny, grid.mg, grid.h
    G = np.zeros((grid.nΓ, mg), order="F")
    for lev in range(1, mg + 1):
        hc = h * 2 ** (lev - 1)
        i = np.arange(1, nx)
        j = np.arange(1, ny)
        x = (i - nx / 2) * hc + (nx / 2) * h - grid.offx
        y = (j - ny / 2) * hc + (ny / 2) * h - grid.offy
        X, Y = np.meshgrid(x, y, indexing="ij")
        r2 = (X - x0) ** 2 + (Y - y0) ** 2
        om = circ / (math.pi * width ** 2) * np.exp(-r2 / width ** 2)
        G[:, lev - 1] = (hc * hc * om).ravel(order="F")
    return G
