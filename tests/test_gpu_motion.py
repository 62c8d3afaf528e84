"""Parity of the lab-frame moving-body path (MotionFunction) on the GPU with the
oracle, through the C ABI.

Per step the library moves the bodies on the host (the oracle's formula, so
xb and body_idx are bit-exact), rebuilds E, and solves the Schur system by
the CG of cg! with one device B_times! per product.  The oracle runs the same
CG over its own B_times!: whole steps agree to the 1e-10 bar of the drop-in.
"""
import numpy as np
import pytest

import ibpm
import oracle as O
from cases import boundary_of, case, gaussian_vortex_gamma, oracle_bodies, oracle_steps, sc7
from conftest import relerr
from ibpm import _capi as C

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-10


def build(c):
    g = c["grid"]
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), motion=2)
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    return op, gp


def initial_states(c, op, gp, seed=11):
    g = c["grid"]
    rng = np.random.default_rng(seed)
    G = gaussian_vortex_gamma(g, x0=0.9, y0=-0.3, circ=0.5, width=0.3)
    G[:, 0] += 1e-3 * g.h * g.h * rng.standard_normal(g.nΓ)
    S = op.new_state(sc7(c, 0.0))
    psi, q, G2 = op.vort2flux(G)
    S.G[:], S.psi[:], S.q[:] = G2, psi, q
    st = ibpm.IBState(gp)
    st.Γ = G
    ibpm.vort2flux_(st, gp)
    return S, st


def test_moving_body_steps_match_oracle():
    c = case("lab")
    op, gp = build(c)
    assert gp.motion_kind == C.MOTION_FUNCTION
    S, st = initial_states(c, op, gp)
    n = 25
    ts = np.arange(n) * c["dt"]
    cd_o, cl_o = op.run(S, oracle_steps(c, ts))
    cd, cl, done = ibpm.advance_n_(st, gp, ts)
    assert done == n
    # the body moved (rotation 0.4 rad/s and translation): indices and points as the oracle's
    xb_o, ub_o = op.body_state()
    xb, ub = gp.body_state()
    assert np.array_equal(xb, xb_o) and np.array_equal(ub, ub_o)
    assert np.array_equal(gp.body_idx, op.body_idx())
    assert not np.array_equal(xb, c["bodies"][0].x0)
    assert np.array_equal(st.xb[0], xb)
    it, res, tol = gp.cg_stats()
    it_o, res_o, tol_o = op.cg_stats()
    assert 0 < it and res <= tol and abs(it - it_o) <= 1
    assert relerr(cd, cd_o) < STEP_TOL
    assert np.max(np.abs(cl - cl_o)) < STEP_TOL * np.abs(cd_o).max()
    errs = {"q": relerr(st.q, S.q), "Γ": relerr(st.Γ, S.G), "ψ": relerr(st.ψ, S.psi), "F̃b": relerr(st.F̃b, S.Fb),
            "nonlin1": relerr(st.nonlin[0], S.nl1)}
    assert all(v < STEP_TOL for v in errs.values()), errs


def test_body_at_rest_matches_static_path():
    """xc = uc = 0: the CG path repeats the dense-Binv Static path."""
    c = case("lab")
    g = c["grid"]
    still = ibpm.MotionFunction([lambda t: 0.0] * 3, [lambda t: 0.0] * 3)
    x0 = c["bodies"][0].x0
    ds = c["bodies"][0].ds
    pm = ibpm.IBProblem(g, [ibpm.RigidBody(still, x0.copy(), x0.copy(), 0 * x0, ds)], c["dt"], c["Re"],
                        freestream=c["freestream"])
    ps = ibpm.IBProblem(g, [ibpm.RigidBody(ibpm.Static(), x0.copy(), x0.copy(), 0 * x0, ds)], c["dt"], c["Re"],
                        freestream=c["freestream"])
    sm, ss = ibpm.IBState(pm), ibpm.IBState(ps)
    ts = np.arange(15) * c["dt"]
    cdm, clm, _ = ibpm.advance_n_(sm, pm, ts)
    cds, cls, _ = ibpm.advance_n_(ss, ps, ts)
    assert relerr(cdm, cds) < 1e-9
    assert np.max(np.abs(clm - cls)) < 1e-9 * np.abs(cds).max()
    assert relerr(sm.Γ, ss.Γ) < 1e-9


def test_update_bodies_and_E_match_oracle(rng):
    c = case("lab")
    op, gp = build(c)
    t = 0.37
    gp.update_bodies(t)
    op.update_bodies(np.array(c["bodies"][0].motion.kinematics(t)))
    assert np.array_equal(gp.body_idx, op.body_idx())
    q = rng.standard_normal(c["grid"].nq)
    fb = np.zeros(gp.nf)
    gp.ctx.check(C.lib().ibpm_op_regT(gp.h, C.ptr(q), C.ptr(fb)))
    assert relerr(fb, op.regT(q)) < 1e-13
    f = rng.standard_normal(gp.nf)
    qq = np.zeros(c["grid"].nq)
    gp.ctx.check(C.lib().ibpm_op_reg(gp.h, C.ptr(f), C.ptr(qq)))
    assert relerr(qq, op.reg(f)) < 1e-13
    with pytest.raises(C.IBPMError):
        gp.Binv  # no matrix: CG every step


def test_leaving_the_grid_is_an_error_and_keeps_the_state():
    c = case("lab")
    g = c["grid"]
    away = ibpm.MotionFunction([lambda t: 3.0 * t, lambda t: 0.0, lambda t: 0.0],
                               [lambda t: 3.0, lambda t: 0.0, lambda t: 0.0])
    x0 = c["bodies"][0].x0
    body = ibpm.RigidBody(away, x0.copy(), x0.copy(), 0 * x0, c["bodies"][0].ds)
    gp = ibpm.IBProblem(g, [body], c["dt"], c["Re"], freestream=c["freestream"])
    st = ibpm.IBState(gp)
    ibpm.advance_(st, gp, 0.1)  # x = 0.3: still inside
    G1 = st.Γ.copy()
    xb1 = body.xb.copy()
    with pytest.raises(C.IBPMError) as e:
        ibpm.advance_(st, gp, 0.7)  # x = 2.1: the window would leave the grid
    assert e.value.code == C.IBPM_E_ARG
    assert np.array_equal(st.Γ, G1)
    assert np.array_equal(body.xb, xb1)


def test_body_leaving_domain_fails_step_and_keeps_problem_usable():
    """A lab-frame body driven towards the domain edge: the step that would
    bring a point within 6 cells of it fails with IBPM_E_ARG (the reference's
    E window raises a BoundsError there), the steps before it are returned,
    and the problem steps on afterwards."""
    import ctypes
    import math

    grid = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -2.0, 2.0), mg=2)
    mot = ibpm.MotionFunction([lambda t: 20.0 * t, lambda t: 0.0, lambda t: 0.0],
                              [lambda t: 20.0, lambda t: 0.0, lambda t: 0.0])
    bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, motion=mot, n=40)]
    gp = ibpm.IBProblem(grid, bodies, 0.01, 100.0, freestream={"Ux": lambda t: 1.0})
    st = ibpm.IBState(gp)
    # x = 20 t: the right edge of the cylinder (x = 0.5 + 20 t) must stay below 2.1875
    # (body_idx + 6 <= nx): steps up to t = 0.08 pass, the step at t = 0.09 fails
    ts = np.arange(20) * 0.01
    arr, keep = gp.step_scalars(ts)
    cd = np.zeros((20, 1))
    cl = np.zeros((20, 1))
    done = ctypes.c_int(-1)
    rc = C.lib().ibpm_advance_n(gp.h, 20, C.ptr(ts), arr, C.ptr(cd), C.ptr(cl), ctypes.byref(done))
    assert rc == C.IBPM_E_ARG
    k = done.value
    assert k == 9 and np.all(np.isfinite(cd[:k]))
    print(f"steps before the failing one: {k}")
    # the failing step left the body where the last finished step put it
    xb, _ = gp.body_state()
    assert math.isclose(float(np.max(xb[:, 0])), 0.5 + 20.0 * ts[k - 1], rel_tol=0, abs_tol=1e-12)
    # the problem is still usable: steps at times the body stays inside
    ts2 = ts[:3]
    cd2, cl2, done2 = ibpm.advance_n_(st, gp, ts2)
    assert done2 == 3 and np.all(np.isfinite(cd2))
