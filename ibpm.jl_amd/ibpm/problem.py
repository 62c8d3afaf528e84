"""IBProblem / IBState / advance! backed by the HIP library.

IBModel, IBProblem     -- src/pre-processing/models.jl:115-158, problem-types.jl:23-41
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
    (problem-types.jl:33)."""
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
    """IBProblem(grid, bodies, dt, Re; freestream) -- problem-types.jl:23-41.

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
            self.motion_kind = C.MOTION_FUNCTION
            if Binv is not None:
                raise ValueError("MotionFunction problems have no Binv matrix (CG every step, laplacian.jl:123-141)")
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

    def set_option(self, option: int, value: int):
        """ibpm_set_option: launch shapes, in-step profiling, graph replay and
        the precision of the Laplacian inverses (include/ibpm_hip.h)."""
        self.ctx.check(C.lib().ibpm_set_option(self.h, int(option), int(value)), "ibpm_set_option")

    def get_option(self, option: int) -> int:
        v = ctypes.c_int(0)
        self.ctx.check(C.lib().ibpm_get_option(self.h, int(option), ctypes.byref(v)), "ibpm_get_option")
        return v.value

    def set_precision(self, precision: str = "fp64", sweeps: int = 2):
        """Arithmetic of the Laplacian inverses in advance!: "fp64" (the
        drop-in), or the mixed-precision form "fp32" / "bf16x3": every DST-I
        direction a dense sine-matrix product on the MFMA units (FP32, or
        split BF16 with FP32 accumulation), applied in increment form to FP64
        residuals, `sweeps` (1 or 2) times per solve."""
        code = {"fp64": C.PREC_FP64, "fp32": C.PREC_FP32_MFMA, "bf16x3": C.PREC_BF16X3_MFMA}[precision]
        self.set_option(C.OPT_PRECISION, code)
        self.set_option(C.OPT_MX_SWEEPS, sweeps)

    @property
    def body_idx(self) -> np.ndarray:
        out = np.zeros((self.nb_total, 2), dtype=np.int32, order="F")
        self.ctx.check(C.lib().ibpm_problem_get_body_idx(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                         out.size))
        return out

    def scalars(self, t: float) -> C.StepScalars:
        """The problem's closures evaluated at time t.  For MotionFunction
        bodies the kinematics buffer stays valid until the next call."""
        arr, keep = self.step_scalars([t])
        self._keep = keep
        return arr[0]

    def step_scalars(self, ts):
        """ibpm_step_scalars for each time in ts, and the buffer that the
        MotionFunction kinematics pointers refer to (keep it alive for the
        call that takes the array)."""
        n = len(ts)
        arr = (C.StepScalars * max(n, 1))()
        fs = self.model.freestream
        kin = None
        if self.motion_kind == C.MOTION_FUNCTION:
            kin = np.zeros((max(n, 1), len(self.model.bodies), 6))
        for k, t in enumerate(ts):
            t = float(t)
            s = arr[k]
            s.Ux, s.Uy, s.inclination = float(fs["Ux"](t)), float(fs["Uy"](t)), float(fs["inclination"](t))
            if self.motion_kind == C.MOTION_MOVING_GRID:
                m = self.model.bodies[0].motion
                s.U, s.V, s.theta, s.theta_dot = float(m.U(t)), float(m.V(t)), float(m.θ(t)), float(m.θ̇(t))
            elif kin is not None:
                for j, b in enumerate(self.model.bodies):
                    kin[k, j, :] = b.motion.kinematics(t)
                s.body_kin = C.ptr(kin[k])
        return arr, kin

    def body_state(self):
        """(xb, ub), each (nb_total, 2): the body points after the last
        update_bodies! and their velocities."""
        xb = np.zeros((self.nb_total, 2), order="F")
        ub = np.zeros((self.nb_total, 2), order="F")
        self.ctx.check(C.lib().ibpm_get_body_state(self.h, C.ptr(xb), C.ptr(ub), xb.size), "ibpm_get_body_state")
        return xb, ub

    def _sync_bodies(self, state=None):
        """update_bodies! mutates bodies[j].xb and .ub (move-body-utils.jl:128-132)
        and sets state.xb[j] = bodies[j].xb (timestepping.jl:392)."""
        if self.motion_kind != C.MOTION_FUNCTION:
            return
        xb, ub = self.body_state()
        k = 0
        for b in self.model.bodies:
            n = len(b.ds)
            b.xb[...] = xb[k:k + n]
            b.ub[...] = ub[k:k + n]
            k += n
        if state is not None:
            state.xb = [b.xb.copy() for b in self.model.bodies]

    def update_bodies(self, t: float):
        """update_bodies!(state, prob, t) alone (timestepping.jl:386-398)."""
        if self.motion_kind != C.MOTION_FUNCTION:
            return
        arr, keep = self.step_scalars([t])
        self.ctx.check(C.lib().ibpm_op_update_bodies(self.h, arr[0].body_kin), "ibpm_op_update_bodies")
        self._sync_bodies()

    def cg_stats(self):
        """(products, final |r|, tolerance) of the last step's CG solve (MotionFunction)."""
        it = ctypes.c_int32(0)
        r = ctypes.c_double(0.0)
        tol = ctypes.c_double(0.0)
        self.ctx.check(C.lib().ibpm_cg_stats(self.h, ctypes.byref(it), ctypes.byref(r), ctypes.byref(tol)))
        return it.value, r.value, tol.value


class LinearizedIBProblem(IBProblem):
    """LinearizedIBProblem(base_state, base_prob, dt) -- problem-types.jl:47-87.

    Same grid, bodies and Re as base_prob (static bodies); prob.Binv is
    base_prob's when dt is the same (get_AB gives the same matrix), else it is
    assembled for dt.  The library keeps QB = avg_flux(base_state, base_prob)
    per level and ΓB = base_state.Γ; nonlinear! forms Q x ΓB + QB x Γ and
    base_flux! sets q0 = 0.  Like the reference, the freestream of the base
    problem is carried along but not used by the step.
    """

    def __init__(self, base_state: "IBState", base_prob: IBProblem, dt: float, device: int = 0):
        m = base_prob.model
        same_dt = float(dt) == float(base_prob.scheme.dt)
        super().__init__(m.grid, m.bodies, dt, m.Re, freestream=m.freestream, device=device,
                         Binv=base_prob.Binv if same_dt else None)
        if self.motion_kind != C.MOTION_STATIC:
            raise NotImplementedError("LinearizedIBProblem is defined for static bodies")
        qB = np.asfortranarray(base_state.q)
        q0B = np.asfortranarray(base_state.q0)
        GB = np.asfortranarray(base_state.Γ)
        self.ctx.check(C.lib().ibpm_problem_linearize(self.h, C.ptr(qB), C.ptr(q0B), C.ptr(GB)),
                       "ibpm_problem_linearize")
        self.base_state = snapshot(base_state)

    def scalars(self, t: float) -> C.StepScalars:
        return C.StepScalars()  # unused by the linearized step (q0 = 0)


_FIELDS = {"q": C.FIELD_Q, "q0": C.FIELD_Q0, "Γ": C.FIELD_GAMMA, "ψ": C.FIELD_PSI, "F̃b": C.FIELD_FB_TILDE,
           "CD": C.FIELD_CD, "CL": C.FIELD_CL}


class IBState:
    """IBState(prob) -- state-types.jl:25-47 (zeros, then base_flux! at t = 0)."""

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
    arr, keep = prob.step_scalars([t])
    try:
        prob.ctx.check(C.lib().ibpm_advance(prob.h, float(t), arr), "ibpm_advance")
    finally:
        state.invalidate()
        prob._sync_bodies(state)
    return None


def advance_n_(state: IBState, prob: IBProblem, ts):
    """advance! at each time of `ts` with no host round trip in between.
    Returns (CD, CL) histories of shape (len(ts), nbodies) and the number of
    steps taken (fewer than len(ts) if a CL became non-finite)."""
    ts = np.asarray(ts, dtype=np.float64)
    n = len(ts)
    arr, keep = prob.step_scalars(ts)
    nbod = max(len(prob.model.bodies), 1)
    cd = np.zeros((n, nbod))
    cl = np.zeros((n, nbod))
    done = ctypes.c_int(0)
    rc = C.lib().ibpm_advance_n(prob.h, n, C.ptr(ts), arr, C.ptr(cd), C.ptr(cl), ctypes.byref(done))
    state.invalidate()
    prob._sync_bodies(state)
    if rc not in (C.IBPM_OK, C.IBPM_E_NONFINITE):
        prob.ctx.check(rc, "ibpm_advance_n")
    return cd[: done.value], cl[: done.value], done.value


def advance_batch_(states, probs, ts_list):
    """advance! of several independent problems at once: each problem's steps
    are enqueued on its own stream (ibpm_advance_async) before any is waited
    for (ibpm_advance_wait), so the GPU runs them concurrently.  Returns one
    (CD, CL, steps_done) per problem, as advance_n_ does."""
    L = C.lib()
    for st, prob, ts in zip(states, probs, ts_list):
        ts = np.asarray(ts, dtype=np.float64)
        arr, keep = prob.step_scalars(ts)
        prob.ctx.check(L.ibpm_advance_async(prob.h, len(ts), arr), "ibpm_advance_async")
    out = []
    for st, prob, ts in zip(states, probs, ts_list):
        n = len(ts)
        nbod = max(len(prob.model.bodies), 1)
        cd = np.zeros((n, nbod))
        cl = np.zeros((n, nbod))
        done = ctypes.c_int(0)
        rc = L.ibpm_advance_wait(prob.h, C.ptr(cd), C.ptr(cl), ctypes.byref(done))
        st.invalidate()
        prob._sync_bodies(st)
        if rc not in (C.IBPM_OK, C.IBPM_E_NONFINITE):
            prob.ctx.check(rc, "ibpm_advance_wait")
        out.append((cd[: done.value], cl[: done.value], done.value))
    return out


class Group:
    """Problems with static bodies on one grid, stepped together: every pass
    of advance! is one launch over all of them (ibpm_group_*).  While the
    group exists the members' states live in its arrays; they are read as
    usual (IBState fields, forces) and stepped only through the group.
    close() (or garbage collection) hands every member its state back."""

    def __init__(self, probs, states=None):
        self.probs = list(probs)
        self.states = list(states) if states is not None else []
        self.ctx = self.probs[0].ctx
        arr = (ctypes.c_void_p * len(self.probs))(*[p.h.value for p in self.probs])
        h = ctypes.c_void_p()
        self.ctx.check(C.lib().ibpm_group_create(self.ctx.h, len(self.probs), arr, ctypes.byref(h)),
                       "ibpm_group_create")
        self.h = h

    def set_option(self, option: int, value: int):
        self.ctx.check(C.lib().ibpm_group_set_option(self.h, option, value), "ibpm_group_set_option")

    def advance_async(self, nsteps: int):
        self.ctx.check(C.lib().ibpm_group_advance_async(self.h, int(nsteps)), "ibpm_group_advance_async")
        self._pending = getattr(self, "_pending", 0) + int(nsteps)

    def wait(self):
        """[(CD, CL, steps_done)] per member for the steps since the last wait."""
        m = getattr(self, "_pending", 0)
        nb = [max(len(p.model.bodies), 1) for p in self.probs]
        cd = np.zeros(max(m * sum(nb), 1))
        cl = np.zeros_like(cd)
        done = (ctypes.c_int32 * len(self.probs))()
        rc = C.lib().ibpm_group_wait(self.h, C.ptr(cd), C.ptr(cl), done)
        self._pending = 0
        for st in self.states:
            st.invalidate()
        if rc not in (C.IBPM_OK, C.IBPM_E_NONFINITE):
            self.ctx.check(rc, "ibpm_group_wait")
        out, o = [], 0
        for i, k in enumerate(nb):
            blk = m * k
            d = int(done[i])
            out.append((cd[o:o + blk].reshape(m, k)[:d], cl[o:o + blk].reshape(m, k)[:d], d))
            o += blk
        return out

    def advance_n_(self, nsteps: int):
        self.advance_async(nsteps)
        return self.wait()

    def close(self):
        h = getattr(self, "h", None)
        if h:
            self.h = None
            rc = C.lib().ibpm_group_destroy(h)
            for st in self.states:
                st.invalidate()
            self.ctx.check(rc, "ibpm_group_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def advance_groups_(groups, nsteps: int):
    """Step several groups together (ibpm_groups_advance_async): one step of
    every group is a single HIP graph with a branch per group, replayed
    nsteps times; returns each group's wait() result."""
    arr = (ctypes.c_void_p * len(groups))(*[g.h.value for g in groups])
    groups[0].ctx.check(C.lib().ibpm_groups_advance_async(len(groups), arr, int(nsteps)),
                        "ibpm_groups_advance_async")
    for g in groups:
        g._pending = getattr(g, "_pending", 0) + int(nsteps)
    return [g.wait() for g in groups]


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


# ---------------------------------------------------------------------------
# save hooks (save_info.jl) and the run_sim! loop (ibpm.jl:90-106)
# ---------------------------------------------------------------------------
class UserVar:
    """user_var (save_info.jl:1-9): values saved at times `t` by `fcn(t, state)`."""

    def __init__(self, val, t, fcn):
        self.val = val
        self.t = np.atleast_1d(np.asarray(t, dtype=np.float64))
        self.fcn = fcn


def init_save_data(t, save_info, state):
    """init_save_data(t, save_info, state) -- save_info.jl:11-94.

    save_info None: one UserVar that keeps the whole state at t[end] (for a
    restart).  Otherwise a dict with "save_fcns" (list of fcn(t, state)) and
    optionally "save_times" (per function: None or 0 -> every step; a float
    -> every round(s/dt) dt from t[0]; a list -> the step times nearest to
    each entry) and "save_types" (ignored: values are kept in lists).  The
    reference's list branch picks argmin(t .- s) and then names an undefined
    variable; this takes the nearest step time, which is what its comment says.
    """
    t = np.asarray(t, dtype=np.float64)
    if save_info is None:
        return [UserVar([], t[-1], lambda tt, st: snapshot(st))]
    dt = t[1] - t[0]
    data = []
    fcns = save_info["save_fcns"]
    times = save_info.get("save_times")
    for j, f in enumerate(fcns):
        if times is None:
            tsv = t[-1:]
        else:
            sv = times[j]
            if sv is None or (np.isscalar(sv) and float(sv) == 0.0):
                tsv = t
            elif np.isscalar(sv):
                step = round(float(sv) / dt) * dt
                n = int(np.floor((t[-1] - t[0]) / step + 1e-9)) + 1
                tsv = t[0] + step * np.arange(n)
            else:
                tsv = np.unique([t[int(np.argmin(np.abs(t - s)))] for s in sv])
        data.append(UserVar([], tsv, f))
    return data


def save_data_(tnow, t, prob, state, data):
    """save_data!(tnow, t, prob, state, data) -- save_info.jl:96-106: append
    fcn(tnow, state) to every UserVar with a save time equal to tnow (to the
    reference's relative tolerance, |ts - tnow| + 1 ≈ 1)."""
    for d in data:
        if np.any(np.isclose(np.abs(d.t - tnow) + 1.0, 1.0, rtol=np.sqrt(np.finfo(float).eps), atol=0.0)):
            d.val.append(d.fcn(tnow, state))


def _save_steps(t, data):
    """Indices of t at which some UserVar saves."""
    if not data:
        return set()
    idx = set()
    tol = np.sqrt(np.finfo(float).eps)
    for d in data:
        for ts in d.t:
            k = np.nonzero(np.isclose(np.abs(t - ts) + 1.0, 1.0, rtol=tol, atol=0.0))[0]
            idx.update(int(i) for i in k)
    return idx


def run_sim_(t, state: IBState, prob: IBProblem, display_freq: int = 25, data=None, verbose: bool = False):
    """run_sim!(t, state, prob; display_freq, data) -- ibpm.jl:90-106.

    Advances over the times t on the device in chunks of steps, each chunk
    ending at the next display or save step: stops after the first step
    whose CL is non-finite; every display_freq steps stores compute_cfl in
    state.cfl; at the save times of `data` (a list of UserVar from
    init_save_data) calls save_data!.  Returns the (CD, CL) histories."""
    t = np.asarray(t, dtype=np.float64)
    saves = _save_steps(t, data or [])
    stops = sorted(saves | {k for k in range(len(t)) if (k + 1) % display_freq == 0} | {len(t) - 1})
    cds, cls = [], []
    i = 0
    for stop in stops:
        if stop < i:
            continue
        chunk = t[i: stop + 1]
        cd, cl, done = advance_n_(state, prob, chunk)
        cds.append(cd)
        cls.append(cl)
        i += done
        if done < len(chunk):
            break
        if (stop + 1) % display_freq == 0:
            state.cfl = compute_cfl(state, prob)
            if verbose:
                print((float(t[stop]), state.CD, state.CL, state.cfl))
        if data and stop in saves:
            save_data_(float(t[stop]), t, prob, state, data)
    cd = np.concatenate(cds) if cds else np.zeros((0, 1))
    cl = np.concatenate(cls) if cls else np.zeros((0, 1))
    return cd, cl


# ---------------------------------------------------------------------------
# restart: a full IBState round trip (state-types.jl:61-98 copy!)
# ---------------------------------------------------------------------------
class StateSnapshot:
    """Host copy of every IBState field (copy! of state-types.jl:62-77):
    q, q0, Γ, ψ, nonlin[1:2], F̃b, fb, CD, CL, xb, cfl, slip."""

    FIELDS = ("q", "q0", "Γ", "ψ", "F̃b", "CD", "CL")

    def __init__(self, **kw):
        self.__dict__.update(kw)


def snapshot(state: IBState) -> StateSnapshot:
    """Host copy of the device state."""
    d = {k: np.array(state._get(k)) for k in StateSnapshot.FIELDS}
    d["nonlin"] = [np.array(a) for a in state.nonlin]
    d["fb"] = [np.array(a) for a in state.fb]
    d["xb"] = [np.array(x) for x in state.xb]
    d["cfl"], d["slip"] = state.cfl, state.slip
    return StateSnapshot(**d)


def copy_(w: IBState, v: StateSnapshot) -> IBState:
    """copy!(w, v) with v a snapshot: uploads every field into w's device
    state.  Continuing from it gives the same steps, bit for bit, as the
    state the snapshot was taken from."""
    # q0 is left alone when it already matches (the device then keeps its
    # uniform-base-flux fast path for static bodies)
    if not np.array_equal(w._get("q0"), v.q0):
        w._set("q0", v.q0)
    for k in ("q", "Γ", "ψ", "F̃b", "CD", "CL"):
        w._set(k, getattr(v, k))
    w.nonlin = v.nonlin
    w.xb = [np.array(x) for x in v.xb]
    w.cfl, w.slip = v.cfl, v.slip
    w.invalidate()
    return w


def save_state(path, state: IBState, t: float = 0.0):
    """Write a restart file (.npz, arrays only: no pickled objects)."""
    s = snapshot(state)
    g = state.prob.model.grid
    arrays = {"q": s.q, "q0": s.q0, "Gamma": s.Γ, "psi": s.ψ, "Fb_tilde": s.F̃b, "CD": s.CD, "CL": s.CL,
              "nonlin1": s.nonlin[0], "nonlin2": s.nonlin[1],
              "meta": np.array([g.nx, g.ny, g.mg, g.h, t, s.cfl, s.slip], dtype=np.float64)}
    for k, x in enumerate(s.xb):
        arrays[f"xb{k}"] = x
    for k, f in enumerate(s.fb):
        arrays[f"fb{k}"] = f
    np.savez(path, **arrays)


def load_state(path, state: IBState) -> float:
    """Read a restart file written by save_state into `state` (whose problem
    must have the same grid); returns the saved time."""
    with np.load(path, allow_pickle=False) as z:
        g = state.prob.model.grid
        nx, ny, mg, h, t, cfl, slip = z["meta"]
        if (int(nx), int(ny), int(mg), float(h)) != (g.nx, g.ny, g.mg, g.h):
            raise ValueError("restart file is for another grid")
        nb = len(state.prob.model.bodies)
        snap = StateSnapshot(q=z["q"], q0=z["q0"], Γ=z["Gamma"], ψ=z["psi"], F̃b=z["Fb_tilde"], CD=z["CD"],
                             CL=z["CL"], nonlin=[z["nonlin1"], z["nonlin2"]],
                             fb=[z[f"fb{k}"] for k in range(nb)], xb=[z[f"xb{k}"] for k in range(nb)],
                             cfl=float(cfl), slip=float(slip))
    copy_(state, snap)
    return float(t)
