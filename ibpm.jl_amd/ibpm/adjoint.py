"""Tangent-linear and adjoint runs of advance! (MovingGrid bodies) and a
torch.autograd.Function over them -- the Python counterpart of the
ChainRulesCore rrule a Julia host would define.

The objective of a run of steps k = 0..n-1 from the problem's current state is

    J = sum_k (wcd[k] CD_k + wcl[k] CL_k),

CD_k, CL_k being state.CD, state.CL after step k (timestepping.jl:330-354).
The library differentiates J with respect to each step's scalars {Ux, Uy,
inclination, U, V, θ, θ̇} (the closures of the problem evaluated at the step's
time, timestepping.jl:284-298) and to the initial Γ; the chain rule through
whatever parameters produce those scalars is the caller's (here: a pitching
motion θ(t) = A sin(2πft) with inflow direction β).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _capi as C

SCALARS = ("Ux", "Uy", "inclination", "U", "V", "theta", "theta_dot")


def _scalars(rows):
    arr = (C.StepScalars * len(rows))()
    for k, r in enumerate(rows):
        s = arr[k]
        s.Ux, s.Uy, s.inclination, s.U, s.V, s.theta, s.theta_dot = (float(v) for v in r)
    return arr


def tangent(prob, rows, ds, wcd, wcl, dGamma0=None):
    """(J, dJ): the objective and its derivative along (ds, dΓ0); rows and ds
    are (n, 7) arrays of step scalars and their perturbations."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    ds = np.ascontiguousarray(ds, dtype=np.float64)
    wcd = np.ascontiguousarray(wcd, dtype=np.float64)
    wcl = np.ascontiguousarray(wcl, dtype=np.float64)
    J, dJ = ctypes.c_double(), ctypes.c_double()
    g0 = None if dGamma0 is None else C.ptr(np.asfortranarray(dGamma0, dtype=np.float64))
    prob.ctx.check(C.lib().ibpm_tangent_n(prob.h, len(rows), _scalars(rows), C.ptr(ds), g0, C.ptr(wcd), C.ptr(wcl),
                                          ctypes.byref(J), ctypes.byref(dJ)), "ibpm_tangent_n")
    return J.value, dJ.value


def adjoint(prob, rows, wcd, wcl, nsnap=32):
    """(J, sbar, Γ0bar, forward_steps): sbar = dJ/d(step scalars) as an (n, 7)
    array, Γ0bar = dJ/dΓ of the initial state (nΓ x mg), by adjoint steps with
    binomial checkpointing in nsnap snapshots."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    wcd = np.ascontiguousarray(wcd, dtype=np.float64)
    wcl = np.ascontiguousarray(wcl, dtype=np.float64)
    n = len(rows)
    J = ctypes.c_double()
    sbar = np.zeros((n, 7))
    g0 = np.zeros((prob.info.nG, prob.info.mg), order="F")
    fwd = ctypes.c_int64()
    prob.ctx.check(C.lib().ibpm_adjoint_n(prob.h, n, _scalars(rows), C.ptr(wcd), C.ptr(wcl), int(nsnap),
                                          ctypes.byref(J), C.ptr(sbar), C.ptr(g0), ctypes.byref(fwd)),
                   "ibpm_adjoint_n")
    return J.value, sbar, g0, fwd.value


def pitch_scalars(ts, A, f, beta=0.0):
    """Step scalars of a plate pitching as θ(t) = A sin(2πft) and moving with
    unit speed along the body-frame direction β (U = cos β, V = sin β), no
    freestream; and their Jacobian d(rows)/d(A, f, β), shape (n, 7, 3)."""
    ts = np.asarray(ts, dtype=np.float64)
    w = 2.0 * math.pi * f
    rows = np.zeros((len(ts), 7))
    jac = np.zeros((len(ts), 7, 3))
    rows[:, 3] = math.cos(beta)
    rows[:, 4] = math.sin(beta)
    rows[:, 5] = A * np.sin(w * ts)
    rows[:, 6] = w * A * np.cos(w * ts)
    jac[:, 3, 2] = -math.sin(beta)
    jac[:, 4, 2] = math.cos(beta)
    jac[:, 5, 0] = np.sin(w * ts)
    jac[:, 5, 1] = A * np.cos(w * ts) * 2.0 * math.pi * ts
    jac[:, 6, 0] = w * np.cos(w * ts)
    jac[:, 6, 1] = 2.0 * math.pi * A * np.cos(w * ts) - w * A * np.sin(w * ts) * 2.0 * math.pi * ts
    return rows, jac


def mean_drag_weights(n, window_start):
    """Weights of the mean of CD over steps window_start .. n-1."""
    wcd = np.zeros(n)
    wcd[window_start:] = 1.0 / (n - window_start)
    return wcd, np.zeros(n)


def _torch_function():
    import torch

    class PitchingMeanDrag(torch.autograd.Function):
        """J(A, f, β) = mean CD over steps window_start..n-1 of a run of the
        pitching problem `prob` from its current state; backward by the
        library's adjoint (binomial checkpointing in nsnap snapshots)."""

        @staticmethod
        def forward(ctx, params, prob, ts, window_start, nsnap):
            A, f, beta = (float(v) for v in params.detach().cpu().double())
            rows, jac = pitch_scalars(ts, A, f, beta)
            wcd, wcl = mean_drag_weights(len(ts), window_start)
            if params.requires_grad:
                J, sbar, _, _ = adjoint(prob, rows, wcd, wcl, nsnap)
                ctx.save_for_backward(torch.as_tensor(np.einsum("kc,kcp->p", sbar, jac), dtype=params.dtype))
            else:
                J, _ = tangent(prob, rows, np.zeros_like(rows), wcd, wcl)
            return torch.as_tensor(J, dtype=params.dtype)

        @staticmethod
        def backward(ctx, gJ):
            (grad,) = ctx.saved_tensors
            return gJ * grad, None, None, None, None

    return PitchingMeanDrag


def pitching_mean_drag(params, prob, ts, window_start, nsnap=32):
    """torch entry point: J(params) with params = tensor([A, f, β])."""
    return _torch_function().apply(params, prob, ts, window_start, nsnap)
