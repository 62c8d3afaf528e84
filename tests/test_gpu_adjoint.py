"""Tangent-linear and adjoint advance! (MovingGrid) through the C ABI.

Checks, each chosen so that a wrong transpose, sign or index fails it:
  * the tangent against central differences of the GPU's own forward steps;
  * the dot-product identity <J'(x) v, w> = <v, J'(x)^T w>: the tangent's dJ
    for perturbations v = (ds, dΓ0) against the adjoint's cotangents
    (sbar, Γ0bar) contracted with v, to 1e-12;
  * the adjoint with 1, 3 and 32 snapshots gives the same gradient (binomial
    checkpointing recomputes bit-identical states);
  * the gradient of the pitching plate's mean drag with respect to (A, f,
    inflow angle) against central differences of the CPU oracle committed in
    tests/golden (20 steps here; the 4000-step configuration in
    test_plate_gradient_full_run).
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import ibpm
from cases import case, gaussian_vortex_gamma
from ibpm import _capi as C

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def scal_array(rows):
    """ibpm_step_scalars array from rows {Ux, Uy, incl, U, V, θ, θ̇}."""
    arr = (C.StepScalars * len(rows))()
    for k, r in enumerate(rows):
        s = arr[k]
        s.Ux, s.Uy, s.inclination, s.U, s.V, s.theta, s.theta_dot = (float(v) for v in r)
    return arr


def tangent(gp, rows, ds, dG0, wcd, wcl):
    J, dJ = ctypes.c_double(), ctypes.c_double()
    ds = np.ascontiguousarray(ds, dtype=np.float64)
    g0 = None if dG0 is None else C.ptr(np.asfortranarray(dG0, dtype=np.float64))
    gp.ctx.check(C.lib().ibpm_tangent_n(gp.h, len(rows), scal_array(rows), C.ptr(ds), g0, C.ptr(wcd), C.ptr(wcl),
                                        ctypes.byref(J), ctypes.byref(dJ)), "ibpm_tangent_n")
    return J.value, dJ.value


def adjoint(gp, rows, wcd, wcl, nsnap):
    J = ctypes.c_double()
    n = len(rows)
    sbar = np.zeros((n, 7))
    g0 = np.zeros((gp.info.nG, gp.info.mg), order="F")
    fwd = ctypes.c_int64()
    gp.ctx.check(C.lib().ibpm_adjoint_n(gp.h, n, scal_array(rows), C.ptr(wcd), C.ptr(wcl), int(nsnap),
                                        ctypes.byref(J), C.ptr(sbar), C.ptr(g0), ctypes.byref(fwd)),
                 "ibpm_adjoint_n")
    return J.value, sbar, g0, fwd.value


def pitch_rows(ts, A, f, beta):
    w = 2 * math.pi * f
    return np.array([[0.0, 0.0, 0.0, math.cos(beta), math.sin(beta), A * math.sin(w * t), w * A * math.cos(w * t)]
                     for t in ts])


def pitch_jac(ts, A, f, beta):
    """d(rows)/d(A, f, beta): (nsteps, 7, 3)."""
    out = np.zeros((len(ts), 7, 3))
    for k, t in enumerate(ts):
        w = 2 * math.pi * f
        out[k, 3, 2] = -math.sin(beta)
        out[k, 4, 2] = math.cos(beta)
        out[k, 5, 0] = math.sin(w * t)
        out[k, 5, 1] = A * math.cos(w * t) * 2 * math.pi * t
        out[k, 6, 0] = w * math.cos(w * t)
        out[k, 6, 1] = 2 * math.pi * A * math.cos(w * t) - w * A * math.sin(w * t) * 2 * math.pi * t
    return out


def moving_problem(name="moving", vortex=True):
    c = case(name)
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    st = ibpm.IBState(gp)
    if vortex:
        st.Γ = gaussian_vortex_gamma(g, x0=0.8, y0=0.2, circ=0.5, width=0.3)
        ibpm.vort2flux_(st, gp)
    return c, g, gp, st


def test_tangent_matches_central_differences():
    c, g, gp, st = moving_problem()
    n = 16
    ts = np.arange(n) * c["dt"]
    A, f = 15 * math.pi / 180, 0.2
    rows = pitch_rows(ts, A, f, 0.0)
    rng = np.random.default_rng(3)
    wcd, wcl = rng.uniform(0.5, 1.5, n), rng.uniform(-1.0, 1.0, n)
    ds = rng.standard_normal((n, 7)) * np.array([0.1, 0.1, 0.05, 0.1, 0.1, 0.05, 0.05])
    J, dJ = tangent(gp, rows, ds, None, wcd, wcl)
    h = 1e-5
    Jp, _ = tangent(gp, rows + h * ds, np.zeros((n, 7)), None, wcd, wcl)
    Jm, _ = tangent(gp, rows - h * ds, np.zeros((n, 7)), None, wcd, wcl)
    fd = (Jp - Jm) / (2 * h)
    print(f"J {J:.12e}  tangent {dJ:.12e}  central difference {fd:.12e}  rel {abs(dJ - fd) / abs(fd):.2e}")
    assert abs(dJ - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("nsnap", [3, 32])
def test_dot_product_identity(nsnap):
    c, g, gp, st = moving_problem()
    n = 12
    ts = np.arange(n) * c["dt"]
    rows = pitch_rows(ts, 15 * math.pi / 180, 0.2, 0.05)
    rng = np.random.default_rng(11)
    wcd, wcl = rng.standard_normal(n), rng.standard_normal(n)
    ds = rng.standard_normal((n, 7))
    dG0 = rng.standard_normal((g.nΓ, g.mg)) * (g.h * 2.0 ** np.arange(g.mg)) ** 2
    J1, dJ = tangent(gp, rows, ds, dG0, wcd, wcl)
    J2, sbar, G0bar, fwd = adjoint(gp, rows, wcd, wcl, nsnap)
    lhs = dJ
    rhs = float(np.sum(sbar * ds) + np.sum(G0bar * dG0))
    print(f"nsnap {nsnap}: forward steps {fwd}; <J'v, w> = {lhs:.15e}, <v, J'^T w> = {rhs:.15e}, "
          f"rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert J1 == J2
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


def test_snapshot_count_does_not_change_the_gradient():
    c, g, gp, st = moving_problem()
    n = 10
    rows = pitch_rows(np.arange(n) * c["dt"], 15 * math.pi / 180, 0.2, 0.0)
    wcd, wcl = np.full(n, 1.0 / n), np.zeros(n)
    ref = None
    counts = []
    for nsnap in (1, 3, 32):
        J, sbar, G0bar, fwd = adjoint(gp, rows, wcd, wcl, nsnap)
        counts.append(fwd)
        if ref is None:
            ref = (sbar, G0bar)
        else:
            assert np.array_equal(sbar, ref[0]) and np.array_equal(G0bar, ref[1])
    print("forward steps with 1, 3, 32 snapshots:", counts)
    assert counts[0] > counts[1] > counts[2] == 2 * n - 1


def test_runs_leave_the_state_as_they_found_it():
    """ibpm_tangent_n and ibpm_adjoint_n run steps on the problem's own state
    and put it back: Γ, ψ, q, q0, nonlin, F̃b, CD and CL read after each call
    equal those before it, bit for bit."""
    c, g, gp, st = moving_problem()
    rows = pitch_rows(np.arange(6) * c["dt"], 15 * math.pi / 180, 0.2, 0.0)
    ibpm.advance_n_(st, gp, np.arange(3) * c["dt"])  # a state with forces and AB2 history
    names = ("Γ", "ψ", "q", "q0", "F̃b", "CD", "CL")
    st.invalidate()
    before = {k: np.array(getattr(st, k), copy=True) for k in names}
    nl_before = [np.array(a, copy=True) for a in st.nonlin]
    wcd, wcl = np.full(6, 1.0), np.zeros(6)
    tangent(gp, rows, np.zeros((6, 7)), None, wcd, wcl)
    adjoint(gp, rows, wcd, wcl, 2)
    st.invalidate()
    for k in names:
        assert np.array_equal(np.asarray(getattr(st, k)), before[k]), k
    for a, b in zip(st.nonlin, nl_before):
        assert np.array_equal(a, b)


def plate_gradient(nsteps, nsnap=32):
    """dJ/d(A, f, beta) by the adjoint and by the tangent on the pitching plate
    (512 x 256 x 4) of tests/golden/make_adjoint_fd.py."""
    spec = json.load(open(os.path.join(GOLDEN, "adjoint_fd_plate.json" if nsteps == 4000 else
                                       f"adjoint_fd_plate_{nsteps}.json")))
    c = case("plate_pitch_512x256x4")
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream={})
    ibpm.IBState(gp)  # impulsive start from rest
    A, f, beta = spec["A"], spec["f"], spec["beta"]
    ts = np.arange(nsteps) * c["dt"]
    rows = pitch_rows(ts, A, f, beta)
    wcd = np.zeros(nsteps)
    wcd[nsteps // 2:] = 1.0 / (nsteps - nsteps // 2)
    wcl = np.zeros(nsteps)
    J, sbar, _, fwd = adjoint(gp, rows, wcd, wcl, nsnap)
    jac = pitch_jac(ts, A, f, beta)
    grad_adj = np.einsum("kc,kcp->p", sbar, jac)
    grad_tlm = np.array([tangent(gp, rows, jac[:, :, p], None, wcd, wcl)[1] for p in range(3)])
    fd = np.array([spec["grad"][k] for k in ("A", "f", "beta")])
    print(f"{nsteps} steps: J {J!r} (oracle {spec['J']!r}); forward steps {fwd}")
    print("  adjoint ", grad_adj, "\n  tangent ", grad_tlm, "\n  oracle FD", fd)
    return J, spec["J"], grad_adj, grad_tlm, fd


def test_plate_gradient_short_run():
    J, Jo, ga, gt, fd = plate_gradient(20)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert np.allclose(ga, gt, rtol=1e-10, atol=1e-12 * np.max(np.abs(gt)))
    assert np.allclose(ga, fd, rtol=1e-5, atol=1e-6 * np.max(np.abs(fd)))


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "adjoint_fd_plate.json")),
                    reason="tests/golden/adjoint_fd_plate.json (oracle finite differences) not generated")
@pytest.mark.timeout(900)
def test_plate_gradient_full_run():
    J, Jo, ga, gt, fd = plate_gradient(4000)
    assert abs(J - Jo) <= 1e-8 * abs(Jo)
    assert np.allclose(ga, gt, rtol=1e-8, atol=1e-10 * np.max(np.abs(gt)))
    assert np.allclose(ga, fd, rtol=1e-4, atol=1e-5 * np.max(np.abs(fd)))


def test_torch_autograd_function():
    """ibpm.adjoint.pitching_mean_drag as a torch.autograd.Function: its
    backward (the adjoint) against torch.autograd.gradcheck's central
    differences of the forward, on the 20-step pitching plate."""
    import torch

    from ibpm.adjoint import pitching_mean_drag

    spec = json.load(open(os.path.join(GOLDEN, "adjoint_fd_plate_20.json")))
    c = case("plate_pitch_512x256x4")
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream={})
    ibpm.IBState(gp)
    ts = np.arange(20) * c["dt"]
    p = torch.tensor([spec["A"], spec["f"], spec["beta"]], dtype=torch.float64, requires_grad=True)
    J = pitching_mean_drag(p, gp, ts, 10)
    J.backward()
    fd = np.array([spec["grad"][k] for k in ("A", "f", "beta")])
    print("autograd", p.grad.numpy(), "oracle FD", fd)
    assert abs(J.item() - spec["J"]) <= 1e-10 * abs(spec["J"])
    assert np.allclose(p.grad.numpy(), fd, rtol=1e-5, atol=1e-6 * np.max(np.abs(fd)))
    assert torch.autograd.gradcheck(lambda q: pitching_mean_drag(q, gp, ts, 10), (p,), eps=1e-6, atol=1e-7,
                                    rtol=1e-5)
