"""Fluid and structure domains: the host-side constructors of the reference.

MultiGrid / make_grid        -- src/fluid-domain/fluid-domain.jl:10-25,44-73
Motion types                 -- src/structure-domain/motion-types.jl:9-50
RigidBody                    -- src/structure-domain/body-types.jl:10-16
make_plate / make_cylinder   -- src/structure-domain/sample-bodies.jl:17-65
make_naca / naca             -- src/structure-domain/sample-bodies.jl:72-134
MotionType / get_body_info   -- src/structure-domain/move-body-utils.jl:15-42
make_body                    -- src/structure-domain/structure-domain.jl:1-144
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np


# --------------------------------------------------------------------------
# grid
# --------------------------------------------------------------------------
@dataclass
class MultiGrid:
    """Nested uniform grids: mg levels of nx x ny cells, level l spacing h 2^(l-1)."""

    nx: int
    ny: int
    nΓ: int
    nq: int
    mg: int
    offx: float
    offy: float
    len: float
    h: float
    LEFT: int
    RIGHT: int
    BOT: int
    TOP: int

    @property
    def nu(self) -> int:
        return self.ny * (self.nx + 1)

    @property
    def nv(self) -> int:
        return self.nx * (self.ny + 1)

    @property
    def nGamma(self) -> int:
        return self.nΓ

    def split_flux(self, q, lev: int = 1):
        """Views (qx (nx+1) x ny, qy nx x (ny+1)) of level `lev` (1-based) of q."""
        q = np.asarray(q)
        col = q[:, lev - 1] if q.ndim == 2 else q
        qx = col[: self.nu].reshape((self.nx + 1, self.ny), order="F")
        qy = col[self.nu:].reshape((self.nx, self.ny + 1), order="F")
        return qx, qy


def _julia_round(x: float) -> int:
    """Julia's round(): nearest, ties to even (numpy/Python round does the same)."""
    return int(round(x))


def make_grid(h: float, boundary, mg: int = 1) -> MultiGrid:
    """make_grid(h, boundary; mg) -- fluid-domain.jl:44-73."""
    x0, x1, y0, y1 = (float(b) for b in boundary)
    offx = -x0
    offy = -y0
    ln = x1 - x0
    ylen = y1 - y0
    nx = _julia_round(ln / h)
    ny = _julia_round(ylen * nx / ln)
    nG = (nx - 1) * (ny - 1)
    nu = ny * (nx + 1)
    nv = nx * (ny + 1)
    hh = ln / nx
    return MultiGrid(nx=nx, ny=ny, nΓ=nG, nq=nu + nv, mg=int(mg), offx=offx, offy=offy, len=ln, h=hh,
                     LEFT=0, RIGHT=ny + 1, BOT=2 * (ny + 1), TOP=2 * (ny + 1) + nx + 1)


# --------------------------------------------------------------------------
# motions
# --------------------------------------------------------------------------
class Motion:
    pass


class InertialMotion(Motion):
    pass


class Static(InertialMotion):
    """Fixed body (motion-types.jl:9-10)."""


@dataclass
class MotionFunction(InertialMotion):
    """Body moving in the lab frame (motion-types.jl:14-17): centre position
    [x, y, θ] and velocity [ẋ, ẏ, θ̇], each a list of three closures of t
    (read as motion.xc[k](t), move-body-utils.jl:77-87).  A single closure
    returning the three values is accepted as well."""

    xc: Any
    uc: Any

    @staticmethod
    def _eval3(f, t):
        if callable(f):
            v = f(t)
        else:
            v = [g(t) for g in f]
        v = [float(x) for x in v]
        if len(v) != 3:
            raise ValueError("MotionFunction xc/uc give three values (x, y, θ) / (ẋ, ẏ, θ̇)")
        return v

    def kinematics(self, t: float):
        """[x, y, θ, ẋ, ẏ, θ̇] at t (get_transformation, move-body-utils.jl:73-97)."""
        return self._eval3(self.xc, t) + self._eval3(self.uc, t)


class MovingGrid(Motion):
    """Body-fixed frame with prescribed U(t), V(t), θ(t), θ̇(t) (motion-types.jl:42-50).

    xc, yc are stored as in the reference, which never reads them: the
    rotation centre is always (0, 0) (models.jl:128, problem-types.jl:36).
    """

    def __init__(self, U: Callable, V: Callable, θ: Callable, θ̇: Callable, xc: float = 0.0, yc: float = 0.0):
        self.U, self.V, self.θ, self.θ̇ = U, V, θ, θ̇
        self.xc, self.yc = float(xc), float(yc)

    @property
    def theta(self):
        return self.θ

    @property
    def theta_dot(self):
        return self.θ̇


@dataclass
class RigidBody:
    """RigidBody{T}(motion, xb, x0, ub, ds) -- body-types.jl:10-16."""

    motion: Motion
    xb: np.ndarray  # (n, 2)
    x0: np.ndarray
    ub: np.ndarray
    ds: np.ndarray  # (n,)


def _body(motion, xb, ds):
    xb = np.ascontiguousarray(xb, dtype=np.float64)
    return RigidBody(motion, xb, xb.copy(), 0.0 * xb, np.asarray(ds, dtype=np.float64))


# --------------------------------------------------------------------------
# sample bodies (sample-bodies.jl)
# --------------------------------------------------------------------------
def make_plate(L, α, h, x0, y0, motion: Motion | None = None, n: int = 0) -> RigidBody:
    """Plate of length L at angle α (sample-bodies.jl:17-33)."""
    if n == 0:
        n = int(math.floor(L / h / 2)) + 1
    k = np.arange(n, dtype=np.float64)
    spt = L * k / (n - 1)
    xhat = spt * math.cos(-α)
    yhat = spt * math.sin(-α)
    xb = np.column_stack([xhat + x0, yhat + y0])
    ds = math.sqrt((xhat[1] - xhat[0]) ** 2 + (yhat[1] - yhat[0]) ** 2)
    return _body(motion or Static(), xb, np.full(n, ds))


def make_cylinder(r, h, x0, y0, motion: Motion | None = None, n: int = 0) -> RigidBody:
    """Circle of radius r (sample-bodies.jl:42-65); n = 0 gives ds ≈ 2h."""
    circum = 2 * math.pi * r
    if n == 0:
        n = int(math.floor(circum / h / 2))
    step = 2 * math.pi / n
    spt = np.arange(n, dtype=np.float64) * step
    xhat = r * np.cos(spt)
    yhat = r * np.sin(spt)
    xb = np.column_stack([xhat + x0, yhat + y0])
    ds = math.sqrt((xhat[1] - xhat[0]) ** 2 + (yhat[1] - yhat[0]) ** 2)
    return _body(motion or Static(), xb, np.full(n, ds))


def naca(x, spec: str):
    """Points of a 4-digit NACA airfoil at chord stations x (sample-bodies.jl:104-134)."""
    x = np.asarray(x, dtype=np.float64)
    m = int(spec[0]) / 100.0
    p = int(spec[1]) / 10.0
    t = int(spec[2:4]) / 100.0
    yc = np.zeros_like(x)
    if m > 0 and p > 0:
        max_idx = np.nonzero(x < p)[0][-1]
        yc[: max_idx + 1] = (m / p ** 2) * (2 * p * x[: max_idx + 1] - x[: max_idx + 1] ** 2)
        yc[max_idx + 1:] = (m / (1 - p) ** 2) * (1 - 2 * p + 2 * p * x[max_idx + 1:] - x[max_idx + 1:] ** 2)
        dyc = np.zeros_like(x)
        dyc[: max_idx + 1] = (2 * m / p ** 2) * (p - x[: max_idx + 1])
        dyc[max_idx + 1:] = (2 * m / (1 - p ** 2)) * (p - x[max_idx + 1:])
        th = np.arctan(dyc)
    else:
        th = np.zeros_like(x)
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)
    xU = x - yt * np.sin(th)
    xL = x + yt * np.sin(th)
    yU = yc + yt * np.cos(th)
    yL = yc - yt * np.cos(th)
    return yc, xU, xL, yU, yL


def make_naca(x0, N: int, spec: str, motion: Motion | None = None) -> RigidBody:
    """NACA airfoil from cosine-spaced panels, points at panel centres (sample-bodies.jl:72-95)."""
    dθ = math.pi / (N - 1)
    x = 0.5 * (1 + np.cos(dθ * np.arange(N)))
    _, xU, xL, yU, yL = naca(x, spec)
    xe = np.zeros((2 * N - 1, 2))
    for i in range(N):
        xe[i] = [x[i] - x0, yU[i]]
    for i in range(1, N):
        xe[i + N - 1] = [x[N - 1 - i] - x0, yL[N - 1 - i]]
    xb = (xe[1:] + xe[:-1]) / 2.0
    ds = np.sqrt(np.sum((xe[1:] - xe[:-1]) ** 2, axis=1))
    return _body(motion or Static(), xb, ds)


# --------------------------------------------------------------------------
# body utilities (move-body-utils.jl)
# --------------------------------------------------------------------------
def MotionType(bodies):
    """Common motion type of the bodies, or Motion if they differ (move-body-utils.jl:15-26)."""
    if isinstance(bodies, RigidBody):
        return type(bodies.motion)
    ts = [type(b.motion) for b in bodies]
    return ts[0] if all(t is ts[0] for t in ts) else Motion


def get_body_info(bodies):
    """(nb, nf): points per body and 2*points per body (move-body-utils.jl:28-33)."""
    nb = [int(b.xb.shape[0]) for b in bodies]
    nf = [2 * n for n in nb]
    return nb, nf


def make_body(bdy_in, Δx):
    """Bodies from NamedTuple-like dicts (structure-domain.jl:1-144)."""
    bodies = []
    for b in bdy_in:
        mo = b.get("motion", "static")
        if mo == "static":
            motion = Static()
        elif mo == "movinggrid":
            f = b.get("motionfcn", {})
            motion = MovingGrid(f.get("U", lambda t: 1.0), f.get("V", lambda t: 0.0), f.get("θ", lambda t: 0.0),
                                f.get("θ̇", lambda t: 0.0), xc=f.get("xc", 0.0), yc=f.get("yc", 0.0))
        elif mo == "movingbody":
            f = b.get("motionfcn", {})
            motion = MotionFunction([f.get("xc", lambda t: -t), f.get("yc", lambda t: 0.0), f.get("θ", lambda t: 0.0)],
                                    [f.get("uc", lambda t: -1.0), f.get("vc", lambda t: 0.0), f.get("θ̇", lambda t: 0.0)])
        else:
            raise ValueError(f"unknown motion {mo}")
        c = b.get("center", (0.0, 0.0))
        if b["type"] == "cylinder":
            bodies.append(make_cylinder(b["lengthscale"], Δx, c[0], c[1], motion=motion))
        elif b["type"] == "plate":
            bodies.append(make_plate(b["lengthscale"], b["spec"]["α"], Δx, c[0], c[1], motion=motion))
        else:
            raise ValueError(f"unknown body type {b['type']}")
    return bodies
