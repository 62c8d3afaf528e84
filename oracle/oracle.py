"""ctypes wrapper of the CPU oracle (``liboracle.so`` built from ``ibpm_oracle.c``).

TEST INFRASTRUCTURE ONLY: the parity tests, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use this module as the checker.  The product
package never imports it.

State arrays use the reference's Julia layouts (column-major, one column per
grid level): ``q, q0`` are ``(nq, mg)``, ``Γ, ψ, nonlin[1:2]`` are ``(nΓ, mg)``,
``F̃b`` is ``(2·nb_total,)`` stacked ``[fx; fy]``, ``fb`` is every body's
``(nb_j, 2)`` block stacked, ``CD, CL`` are per body (src/pre-processing/
state-types.jl:25-47).  Arrays are held as Fortran-ordered numpy arrays so that
``arr.ravel(order='F')`` is the reference's memory image.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_problem_create.restype = ctypes.c_void_p
        L.oracle_problem_create.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                                     ctypes.c_int, _ip, _dp, _dp, ctypes.c_int, _dp, _ip]
        L.oracle_problem_free.argtypes = [ctypes.c_void_p]
        L.oracle_problem_sizes.argtypes = [ctypes.c_void_p, _ip, _dp]
        for name in ("oracle_get_binv", "oracle_get_weights", "oracle_get_lam"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _dp]
        L.oracle_get_body_idx.argtypes = [ctypes.c_void_p, _ip]
        L.oracle_base_flux.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp]
        L.oracle_advance.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _dp]
        L.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_dst_variant.argtypes = [ctypes.c_int]
        L.oracle_delta_h.restype = ctypes.c_double
        L.oracle_delta_h.argtypes = [ctypes.c_double] * 3
        L.oracle_dst2_raw.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_op_dst2.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_op_lapinv.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int]
        for name in ("oracle_op_rot", "oracle_op_curl_sg", "oracle_op_get_bc", "oracle_op_coarsify",
                     "oracle_op_reg", "oracle_op_regT", "oracle_op_B_times"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_op_curl_mg.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.oracle_op_apply_bc.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double]
        L.oracle_op_vort2flux.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.oracle_op_nonlinear.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_invert.argtypes = [_dp, ctypes.c_int]
        L.oracle_problem_linearize.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.oracle_step_stride.argtypes = [ctypes.c_void_p]
        L.oracle_get_body_state.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_cg_stats.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.oracle_op_update_bodies.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_op_cg.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_int]
        L.oracle_invert.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous), "need contiguous float64"
    return a.ctypes.data_as(_dp)


def set_threads(n: int):
    lib().oracle_set_threads(int(n))


def set_dst_variant(v: int):
    """0: the oracle's own 2D DST-I; 1: the same transform in another
    operation order (one line per FFT, dimension 2 first); 2: the HIP
    library's half-length algorithm; 3: that algorithm with a compensated
    sum of its odd outputs."""
    lib().oracle_set_dst_variant(int(v))


def delta_h(rf, rb, dr):
    """δh (src/interface-coupling/interface-coupling.jl:6-62)."""
    return lib().oracle_delta_h(float(rf), float(rb), float(dr))


def dst2(x):
    """2D FFTW RODFT00 of an (n1, n2) array (src/fluid-operators/dst-inversion.jl:10-20)."""
    x = np.asfortranarray(x, dtype=np.float64)
    y = np.zeros_like(x, order="F")
    lib().oracle_dst2_raw(_p(y), _p(x), x.shape[0], x.shape[1])
    return y


def invert(A):
    A = np.array(A, dtype=np.float64, order="F")
    if lib().oracle_invert(_p(A), A.shape[0]) != 0:
        raise np.linalg.LinAlgError("singular")
    return A


class _CState(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("q", "q0", "G", "psi", "nl1", "nl2", "Fb", "fb", "CD", "CL")]


class OState:
    """IBState fields as Fortran-ordered float64 arrays (state-types.jl:25-47)."""

    def __init__(self, P: "OracleProblem"):
        nq, nG, mg, nbt, nbod = P.nq, P.nG, P.mg, P.nbt, P.nbodies
        self.q = np.zeros((nq, mg), order="F")
        self.q0 = np.zeros((nq, mg), order="F")
        self.G = np.zeros((nG, mg), order="F")
        self.psi = np.zeros((nG, mg), order="F")
        self.nl1 = np.zeros((nG, mg), order="F")
        self.nl2 = np.zeros((nG, mg), order="F")
        self.Fb = np.zeros(2 * nbt)
        self.fb = np.zeros(2 * nbt)
        self.CD = np.zeros(max(nbod, 1))
        self.CL = np.zeros(max(nbod, 1))

    def cstruct(self):
        s = _CState()
        for n in ("q", "q0", "G", "psi", "nl1", "nl2", "Fb", "fb", "CD", "CL"):
            setattr(s, n, _p(getattr(self, n)))
        return s

    def copy(self):
        c = object.__new__(OState)
        for n in ("q", "q0", "G", "psi", "nl1", "nl2", "Fb", "fb", "CD", "CL"):
            setattr(c, n, getattr(self, n).copy(order="F"))
        return c


class OracleProblem:
    """IBProblem + IBModel restated on the CPU (problem-types.jl:23-41).

    bodies: list of (xb (n,2), ds (n,)) arrays.  motion: 0 Static, 1 MovingGrid,
    2 MotionFunction (lab-frame moving bodies; xb are the reference points x0).
    binv: optional precomputed inverse Schur matrix (nf x nf); computed by
    get_Binv (laplacian.jl:98-119) when None (no matrix for MotionFunction:
    laplacian.jl:123-141 solves by CG every step).

    Step scalars: 7 per step (Ux, Uy, inclination, U, V, θ, θ̇); for
    MotionFunction followed by {x, y, θ, ẋ, ẏ, θ̇} = motion.xc(t), motion.uc(t)
    of every body (``step_stride`` values in all).
    """

    def __init__(self, h, boundary, mg, Re, dt, bodies, motion=0, binv=None):
        L = lib()
        nb = np.array([len(b[1]) for b in bodies], dtype=np.int32)
        xb = np.concatenate([np.asarray(b[0], dtype=np.float64) for b in bodies], axis=0) if bodies else np.zeros((0, 2))
        xbF = np.asfortranarray(xb)
        ds = np.concatenate([np.asarray(b[1], dtype=np.float64) for b in bodies]) if bodies else np.zeros(0)
        nf = 2 * int(nb.sum())
        bptr = None
        if binv is not None:
            binv = np.asfortranarray(binv, dtype=np.float64)
            assert binv.shape == (nf, nf)
            bptr = _p(binv)
        st = ctypes.c_int(0)
        x0, x1, y0, y1 = boundary
        self._h = L.oracle_problem_create(float(h), float(x0), float(x1), float(y0), float(y1), int(mg), float(Re),
                                          float(dt), len(bodies), nb.ctypes.data_as(_ip), _p(xbF), _p(ds),
                                          int(motion), bptr, ctypes.byref(st))
        if not self._h:
            raise ValueError(f"oracle_problem_create failed with status {st.value}")
        iv = np.zeros(7, dtype=np.int32)
        dv = np.zeros(4)
        L.oracle_problem_sizes(self._h, iv.ctypes.data_as(_ip), _p(dv))
        self.nx, self.ny, self.nG, self.nq, self.mg, self.nbt, self.nbc = (int(v) for v in iv)
        self.h, self.offx, self.offy, self.len = (float(v) for v in dv)
        self.nu = self.ny * (self.nx + 1)
        self.nbodies = len(bodies)
        self.nb = nb
        self.Re, self.dt, self.motion = float(Re), float(dt), int(motion)
        self.step_stride = int(L.oracle_step_stride(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_problem_free(self._h)
            self._h = None

    # --- setup products -------------------------------------------------
    def binv(self):
        nf = 2 * self.nbt
        out = np.zeros((nf, nf), order="F")
        lib().oracle_get_binv(self._h, _p(out))
        return out

    def body_idx(self):
        out = np.zeros((self.nbt, 2), dtype=np.int32, order="F")
        lib().oracle_get_body_idx(self._h, out.ctypes.data_as(_ip))
        return out

    def weights(self):
        """(nbt, 2, 13, 13) array: weight[k, c, l, m] of setup_reg."""
        raw = np.zeros(2 * self.nbt * 169)
        lib().oracle_get_weights(self._h, _p(raw))
        return raw.reshape(self.nbt, 2, 13, 13).transpose(0, 1, 3, 2).copy()

    def lam(self):
        out = np.zeros((self.nx - 1, self.ny - 1), order="F")
        lib().oracle_get_lam(self._h, _p(out))
        return out

    def linearize(self, qB, q0B, GB):
        """Turn this problem into LinearizedIBProblem(base_state, base_prob, dt)
        (problem-types.jl:66-86) around the base state (q, q0, Γ)."""
        qB = np.asfortranarray(qB, dtype=np.float64)
        q0B = np.asfortranarray(q0B, dtype=np.float64)
        GB = np.asfortranarray(GB, dtype=np.float64)
        if lib().oracle_problem_linearize(self._h, _p(qB), _p(q0B), _p(GB)) != 0:
            raise ValueError("LinearizedIBProblem needs static bodies")

    # --- state and stepping ----------------------------------------------
    def new_state(self, sc0):
        """IBState(prob): zeros, then base_flux! at t = 0 (state-types.jl:25-47)."""
        S = OState(self)
        self.base_flux(S, sc0)
        return S

    def base_flux(self, S, sc):
        sc = np.ascontiguousarray(sc, dtype=np.float64)
        cs = S.cstruct()
        lib().oracle_base_flux(self._h, ctypes.byref(cs), _p(sc))

    def advance(self, S, sc):
        sc = np.ascontiguousarray(sc, dtype=np.float64)
        assert sc.size == self.step_stride
        cs = S.cstruct()
        rc = lib().oracle_advance(self._h, ctypes.byref(cs), _p(sc))
        if rc != 0:
            raise ValueError(f"oracle_advance failed with status {rc}")

    def run(self, S, sc_steps, record=True):
        sc_steps = np.ascontiguousarray(sc_steps, dtype=np.float64).reshape(-1, self.step_stride)
        n = sc_steps.shape[0]
        cd = np.zeros((n, max(self.nbodies, 1)))
        cl = np.zeros((n, max(self.nbodies, 1)))
        cs = S.cstruct()
        rc = lib().oracle_run(self._h, ctypes.byref(cs), n, _p(sc_steps), _p(cd) if record else None,
                              _p(cl) if record else None)
        if rc != 0:
            raise ValueError(f"oracle_run failed with status {rc}")
        return cd, cl

    # --- MotionFunction ----------------------------------------------------
    def body_state(self):
        """(xb, ub), each (nbt, 2): the moved points and their velocities."""
        xb = np.zeros((self.nbt, 2), order="F")
        ub = np.zeros((self.nbt, 2), order="F")
        lib().oracle_get_body_state(self._h, _p(xb), _p(ub))
        return xb, ub

    def cg_stats(self):
        """(iterations, final residual norm, tolerance) of the last CG solve."""
        it = ctypes.c_int(0)
        rt = np.zeros(2)
        lib().oracle_cg_stats(self._h, ctypes.byref(it), _p(rt))
        return it.value, float(rt[0]), float(rt[1])

    def update_bodies(self, kin):
        """update_bodies! alone: move the bodies to kin (6 per body) and rebuild E."""
        kin = np.ascontiguousarray(kin, dtype=np.float64)
        rc = lib().oracle_op_update_bodies(self._h, _p(kin))
        if rc != 0:
            raise ValueError(f"update_bodies failed with status {rc}")

    def cg(self, b, x0=None, reltol=1e-12, maxiter=5000):
        """cg!(x, B, b) from the initial guess x0 (zeros when None)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
        lib().oracle_op_cg(self._h, _p(x), _p(b), float(reltol), int(maxiter))
        return x

    # --- operators ---------------------------------------------------------
    def _g(self, n):
        return np.zeros(n)

    def dst2(self, x):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F"))
        y = np.zeros_like(x)
        lib().oracle_op_dst2(self._h, _p(y), _p(x))
        return y

    def lapinv(self, b, kind, lev=1):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().oracle_op_lapinv(self._h, _p(x), _p(b), int(kind), int(lev))
        return x

    def rot(self, q):
        G = np.zeros(self.nG)
        lib().oracle_op_rot(self._h, _p(G), _p(np.ascontiguousarray(q, dtype=np.float64)))
        return G

    def curl_sg(self, psi):
        q = np.zeros(self.nq)
        lib().oracle_op_curl_sg(self._h, _p(q), _p(np.ascontiguousarray(psi, dtype=np.float64)))
        return q

    def curl_mg(self, psi, bc):
        q = np.zeros(self.nq)
        lib().oracle_op_curl_mg(self._h, _p(q), _p(np.ascontiguousarray(psi, dtype=np.float64)),
                                _p(np.ascontiguousarray(bc, dtype=np.float64)))
        return q

    def get_bc(self, r):
        bc = np.zeros(self.nbc)
        lib().oracle_op_get_bc(self._h, _p(bc), _p(np.ascontiguousarray(r, dtype=np.float64)))
        return bc

    def apply_bc(self, r, bc, fac):
        r = np.array(r, dtype=np.float64)
        lib().oracle_op_apply_bc(self._h, _p(r), _p(np.ascontiguousarray(bc, dtype=np.float64)), float(fac))
        return r

    def coarsify(self, Gc, G):
        Gc = np.array(Gc, dtype=np.float64)
        lib().oracle_op_coarsify(self._h, _p(Gc), _p(np.ascontiguousarray(G, dtype=np.float64)))
        return Gc

    def reg(self, fb):
        q = np.zeros(self.nq)
        lib().oracle_op_reg(self._h, _p(q), _p(np.ascontiguousarray(fb, dtype=np.float64)))
        return q

    def regT(self, q):
        fb = np.zeros(2 * self.nbt)
        lib().oracle_op_regT(self._h, _p(fb), _p(np.ascontiguousarray(q, dtype=np.float64)))
        return fb

    def vort2flux(self, G):
        """Returns (ψ, q, Γ_after) for Γ of shape (nΓ, mg)."""
        G = np.array(G, dtype=np.float64, order="F")
        psi = np.zeros((self.nG, self.mg), order="F")
        q = np.zeros((self.nq, self.mg), order="F")
        lib().oracle_op_vort2flux(self._h, _p(psi), _p(q), _p(G))
        return psi, q, G

    def nonlinear(self, q, q0, G, lev):
        out = np.zeros(self.nG)
        q = np.asfortranarray(q, dtype=np.float64)
        q0 = np.asfortranarray(q0, dtype=np.float64)
        G = np.asfortranarray(G, dtype=np.float64)
        lib().oracle_op_nonlinear(self._h, _p(out), _p(q), _p(q0), _p(G), int(lev))
        return out

    def B_times(self, z):
        x = np.zeros(2 * self.nbt)
        lib().oracle_op_B_times(self._h, _p(x), _p(np.ascontiguousarray(z, dtype=np.float64)))
        return x
