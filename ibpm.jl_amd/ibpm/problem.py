"""IBProblem / IBState / advance! backed by the HIP library.

IBModel, IBProblem     -- src/pre-processing/models.jl:115-158, problem-types.jl:182-200
AdamsBashforth, AB2    -- src/pre-processing/models.jl:60-72
IBState                -- src/pre-processing/state-types.jl:7-47
advance!, advance      -- src/timestepping/timestepping.jl:9-67
run_sim!, compute_cfl  -- src/ibpm.jl:83-106

The device holds the authoritative state.  IBState's fields read like the
reference's plain arrays: reading one downloads it (cached until the next
step), assigning one uploads it.  In-place edits of a downloaded array do not
reach the device: assign the array back (``state.Γ = arr``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _capi as C
from .domain import MotionFunction, MovingGrid, MotionType, MultiGrid, RigidBody, Static, get_body_info


@dataclass
class AdamsBashforth:
    dt: float
    β: list = field(default_factory=lambda: [1.5, -0.5])


def AB2(dt: float) -> AdamsBashforth:
    """Second-order Adams-Bashforth (models.jl:70-72)."""
    return AdamsBashforth(float(dt), [1.5, -0.5])


def _zero(t):
    return 0.0


def _freestream(fs) -> dict:
    """Normalise the freestream closures; the reference default is all zero
    (problem-types.jl:192)."""
    fs = dict(fs or {})
    out = {}
    for k in ("Ux", "Uy", "inclination"):
        v = fs.get(k, _zero)
        out[k] = v if callable(v) else (lambda t, c=float(v): c)
    return out


class IBModel:
    """Grid, bodies, Re and freestream of a problem (models.jl:115-158)."""

    def __init__(self, grid: MultiGrid, bodies, Re, freestream=None):
        self.grid = grid
        self.bodies = list(bodies)
        self.Re = float(Re)
        self.freestream = _freestream(freestream)


class IBProblem:
    """IBProblem(grid, bodies, dt, Re; freestream) -- problem-types.jl:188-200.

    The operators (Λ, DST plans, E, Ainv/A per level) and prob.Binv are built by
    the library on `device`; `Binv` may be given to skip the GPU assembly of B.
    """

    def __init__(self, grid: MultiGrid, bodies, dt: float, Re: float, freestream=None, device: int = 0,
                 Binv: np.ndarray | None = None):
        self.model = IBModel(grid, bodies, Re, freestream)
        self.scheme = AB2(dt)
        mt = MotionType(self.model.bodies) if self.model.bodies else Static
        if mt is Static:
            self.motion_kind = C.MOTION_STATIC
        elif mt is MovingGrid:
            self.motion_kind = C.MOTION_MOVING_GRID
            if len(self.model.bodies) != 1:
                raise ValueError("MovingGrid needs exactly one body (timestepping.jl:277)")
        elif mt is MotionFunction:
            raise NotImplementedError("MotionFunction (lab-frame moving body) is outside the GPU path")
        else:
            raise NotImplementedError("bodies with different motion types")
        self.ctx = C.Context.get(device)
        L = C.lib()
        g = grid
        gd = C.GridDesc(nx=g.nx, ny=g.ny, mg=g.mg, h=g.h, offx=g.offx, offy=g.offy, Re=float(Re), dt=float(dt))
        gd.beta[0], gd.beta[1] = self.scheme.β
        nb, _ = get_body_info(self.model.bodies)
        self._nb = np.asarray(nb, dtype=np.int32)
        xb = (np.concatenate([b.xb for b in self.model.bodies], axis=0) if self.model.bodies
              else np.zeros((0, 2)))
        self._xb = np.asfortranarray(xb, dtype=np.float64)
        self._ds = np.ascontiguousarray(np.concatenate([b.ds for b in self.model.bodies])
                                        if self.model.bodies else np.zeros(0), dtype=np.float64)
        bd = C.BodyDesc(nbodies=len(nb), nb=self._nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                        xb=C.ptr(self._xb) if xb.size else None, ds=C.ptr(self._ds) if xb.size else None,
                        motion=self.motion_kind)
        bptr = None
        if Binv is not None:
            self._binv_in = np.asfortranarray(Binv, dtype=np.float64)
            bptr = C.ptr(self._binv_in)
        h = ctypes.c_void_p()
        self.ctx.check(L.ibpm_problem_create(self.ctx.h, ctypes.byref(gd), ctypes.byref(bd), bptr, ctypes.byref(h)),
                       "ibpm_problem_create")
        self.h = h
        info = C.ProblemInfo()
        self.ctx.check(L.ibpm_problem_get_info(self.h, ctypes.byref(info)))
        self.info = info
        self.nb_total = int(info.nb_total)
        self.nf = 2 * self.nb_total

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            C.lib().ibpm_problem_destroy(h)
            self.h = None

    @property
    def Binv(self) -> np.ndarray:
        out = np.zeros((self.nf, self.nf), order="F")
        self.ctx.check(C.lib().ibpm_problem_get_binv(self.h, C.ptr(out), out.size))
        return out

    @property
    def body_idx(self) -> np.ndarray:
        out = np.zeros((self.nb_total, 2), dtype=np.int32, order="F")
        self.ctx.check(C.lib().ibpm_problem_get_body_idx(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                         out.size))
        return out

    def scalars(self, t: float) -> C.StepScalars:
        """The problem's closures evaluated at time t."""
        fs = self.model.freestream
        s = C.StepScalars(Ux=float(fs["Ux"](t)), Uy=float(fs["Uy"](t)), inclination=float(fs["inclination"](t)))
        if self.motion_kind == C.MOTION_MOVING_GRID:
            m = self.model.bodies[0].motion
            s.U, s.V, s.theta, s.theta_dot = float(m.U(t)), float(m.V(t)), float(m.θ(t)), float(m.θ̇(t))
        return s


_FIELDS = {"q": C.FIELD_Q, "q0": C.FIELD_Q0, "Γ": C.FIELD_GAMMA, "ψ": C.FIELD_PSI, "F̃b": C.FIELD_FB_TILDE,
           "CD": C.FIELD_CD, "CL": C.FIELD_CL}


class IBState:
    """IBState(prob) -- state-types.jl:270-293 (zeros, then base_flux! at t = 0)."""

    def __init__(self, prob: IBProblem):
        self.prob = prob
        g = prob.model.grid
        self._shape = {"q": (g.nq, g.mg), "q0": (g.nq, g.mg), "Γ": (g.nΓ, g.mg), "ψ": (g.nΓ, g.mg),
                       "F̃b": (prob.nf, 1), "CD": (len(prob.model.bodies),), "CL": (len(prob.model.bodies),)}
        self._cache: dict = {}
        self.cfl = 0.0
        self.slip = 0.0
        self.xb = [b.xb.copy() for b in prob.model.bodies]
        s = prob.scalars(0.0)
        prob.ctx.check(C.lib().ibpm_state_init(prob.h, ctypes.byref(s)), "ibpm_state_init")

    # --- field access ------------------------------------------------------
    def _get(self, name):
        if name not in self._cache:
            shp = self._shape[name]
            a = np.zeros(shp, order="F")
            self.prob.ctx.check(C.lib().ibpm_state_download(self.prob.h, _FIELDS[name], C.ptr(a), a.size), name)
            self._cache[name] = a
        return self._cache[name]

    def _set(self, name, value):
        a = np.asfortranarray(np.asarray(value, dtype=np.float64).reshape(self._shape[name], order="F"))
        self.prob.ctx.check(C.lib().ibpm_state_upload(self.prob.h, _FIELDS[name], C.ptr(a), a.size), name)
        self._cache.pop(name, None)

    def invalidate(self):
        self._cache.clear()

    q = property(lambda s: s._get("q"), lambda s, v: s._set("q", v))
    q0 = property(lambda s: s._get("q0"), lambda s, v: s._set("q0", v))
    Γ = property(lambda s: s._get("Γ"), lambda s, v: s._set("Γ", v))
    ψ = property(lambda s: s._get("ψ"), lambda s, v: s._set("ψ", v))
    F̃b = property(lambda s: s._get("F̃b"), lambda s, v: s._set("F̃b", v))
    CD = property(lambda s: s._get("CD"), lambda s, v: s._set("CD", v))
    CL = property(lambda s: s._get("CL"), lambda s, v: s._set("CL", v))
    Gamma = Γ
    psi = ψ
    Fb = F̃b

    @property
    def nonlin(self):
        g = self.prob.model.grid
        out = []
        for f in (C.FIELD_NONLIN1, C.FIELD_NONLIN2):
            a = np.zeros((g.nΓ, g.mg), order="F")
            self.prob.ctx.check(C.lib().ibpm_state_download(self.prob.h, f, C.ptr(a), a.size), "nonlin")
            out.append(a)
        return out

    @nonlin.setter
    def nonlin(self, arrs):
        for f, v in zip((C.FIELD_NONLIN1, C.FIELD_NONLIN2), arrs):
            a = np.asfortranarray(np.asarray(v, dtype=np.float64))
            self.prob.ctx.check(C.lib().ibpm_state_upload(self.prob.h, f, C.ptr(a), a.size), "nonlin")

    @property
    def fb(self):
        """Per-body (nb_j, 2) surface stresses."""
        raw = np.zeros(self.prob.nf)
        self.prob.ctx.check(C.lib().ibpm_state_download(self.prob.h, C.FIELD_FB, C.ptr(raw), raw.size), "fb")
        nb, _ = get_body_info(self.prob.model.bodies)
        out, o = [], 0
        for n in nb:
            out.append(raw[2 * o: 2 * o + 2 * n].reshape((n, 2), order="F"))
            o += n
        return out


def base_flux_(state: IBState, prob: IBProblem, t: float):
    """base_flux!(state, prob, t) -- timestepping.jl:230-318."""
    s = prob.scalars(t)
    prob.ctx.check(C.lib().ibpm_base_flux(prob.h, ctypes.byref(s)), "ibpm_base_flux")
    state.invalidate()


def vort2flux_(state: IBState, prob: IBProblem):
    """vort2flux!(state.ψ, state.q, state.Γ, model, mg) on the device state (curl.jl:133-165)."""
    prob.ctx.check(C.lib().ibpm_vort2flux(prob.h), "ibpm_vort2flux")
    state.invalidate()


def advance_(state: IBState, prob: IBProblem, t: float):
    """advance!(state, prob, t) -- timestepping.jl:35-67.  Mutates `state`."""
    s = prob.scalars(t)
    prob.ctx.check(C.lib().ibpm_advance(prob.h, float(t), ctypes.byref(s)), "ibpm_advance")
    state.invalidate()
    return None


def advance_n_(state: IBState, prob: IBProblem, ts):
    """advance! at each time of `ts` with no host round trip in between.
    Returns (CD, CL) histories of shape (len(ts), nbodies) and the number of
    steps taken (fewer than len(ts) if a CL became non-finite)."""
    ts = np.asarray(ts, dtype=np.float64)
    n = len(ts)
    arr = (C.StepScalars * max(n, 1))()
    for k, t in enumerate(ts):
        arr[k] = prob.scalars(float(t))
    nbod = max(len(prob.model.bodies), 1)
    cd = np.zeros((n, nbod))
    cl = np.zeros((n, nbod))
    done = ctypes.c_int(0)
    rc = C.lib().ibpm_advance_n(prob.h, n, C.ptr(ts), arr, C.ptr(cd), C.ptr(cl), ctypes.byref(done))
    state.invalidate()
    if rc not in (C.IBPM_OK, C.IBPM_E_NONFINITE):
        prob.ctx.check(rc, "ibpm_advance_n")
    return cd[: done.value], cl[: done.value], done.value


def advance(x: IBState, prob: IBProblem, t: float):
    """advance(x, prob, t) -- timestepping.jl:22-28.  The device state is
    single-copy per problem, so this advances `x` in place and returns it."""
    advance_(x, prob, t)
    return x


def compute_cfl(state: IBState, prob: IBProblem) -> float:
    """max |q[:,1]| dt / h (ibpm.jl:83-88)."""
    v = ctypes.c_double(0.0)
    prob.ctx.check(C.lib().ibpm_compute_cfl(prob.h, ctypes.byref(v)), "ibpm_compute_cfl")
    return v.value


def run_sim_(t, state: IBState, prob: IBProblem, display_freq: int = 25, data=None, verbose: bool = False):
    """run_sim!(t, state, prob) -- ibpm.jl:90-106: advance over the times t,
    stop after a step whose CL is non-finite, compute the CFL every
    display_freq steps.  `data` may be a callable (t, state) -> None called
    after each chunk of display_freq steps.  Returns (CD, CL) histories."""
    t = np.asarray(t, dtype=np.float64)
    cds, cls = [], []
    i = 0
    while i < len(t):
        chunk = t[i: i + display_freq]
        cd, cl, done = advance_n_(state, prob, chunk)
        cds.append(cd)
        cls.append(cl)
        i += done
        if done < len(chunk):
            break
        if len(chunk) == display_freq:
            state.cfl = compute_cfl(state, prob)
            if verbose:
                print((float(chunk[-1]), state.CD, state.CL, state.cfl))
        if data is not None:
            data(float(chunk[-1]), state)
    cd = np.concatenate(cds) if cds else np.zeros((0, 1))
    cl = np.concatenate(cls) if cls else np.zeros((0, 1))
    return cd, cl
