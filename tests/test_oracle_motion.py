"""Pins of the oracle's lab-frame moving-body path (MotionFunction), CPU only.

MotionFunction (motion-types.jl:14-17) moves the body through a fixed grid:
every step update_bodies! (timestepping.jl:386-398) moves the points by the
rigid transform of move_body! (move-body-utils.jl:73-97,123-133) and rebuilds
E; boundary_forces! (timestepping.jl:176-186) subtracts the point velocities
from E(qs + q0) and solves with B by CG (laplacian.jl:123-141, IterativeSolvers
0.9.1 cg!).  The reference ships no vectors for this path apart from the
end-to-end lift history force_jl1.mat (checked on the GPU), so it is pinned
here by what the mathematics fixes:
  * move_body! against rigid motion written with complex numbers;
  * E after a move against E of a static body set up at the moved points;
  * the CG solve against a dense solve of the assembled B, and the warm start;
  * a body that never moves against the Static path (dense Binv);
  * the first step from rest of a body translating at -U against a static
    body in a freestream U, and of a body spinning at ω against the
    MovingGrid frame spinning at ω: both right-hand sides are the same up to
    the δh moment errors, which pins the sign and scale of both parts of ub.
"""
import numpy as np
import pytest

import ibpm
import oracle as O
from cases import boundary_of, case, oracle_bodies
from conftest import relerr


def _mk(c, motion, bodies=None, freestream=True):
    g = c["grid"]
    return O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], bodies or oracle_bodies(c), motion=motion)


def _sc(P, U=0.0, kin=None, mg=None):
    s = np.zeros(P.step_stride)
    s[0] = U
    if mg is not None:
        s[3:7] = mg
    if kin is not None:
        s[7:] = np.ravel(kin)
    return s


@pytest.fixture(scope="module")
def small():
    return case("small3")


def test_move_body_is_a_rigid_motion(small):
    P = _mk(small, 2)
    x0 = np.asarray(small["bodies"][0].xb)
    X, Y, th, U, V, om = 0.11, -0.07, 0.3, 0.7, -0.2, 1.3
    P.update_bodies([X, Y, th, U, V, om])
    xb, ub = P.body_state()
    z0 = x0[:, 0] + 1j * x0[:, 1]
    z = (X + 1j * Y) + np.exp(1j * th) * z0
    u = (U + 1j * V) + 1j * om * np.exp(1j * th) * z0
    assert np.max(np.abs(xb[:, 0] + 1j * xb[:, 1] - z)) < 1e-14
    assert np.max(np.abs(ub[:, 0] + 1j * ub[:, 1] - u)) < 1e-14


def test_moved_E_is_E_of_the_moved_points(small):
    P = _mk(small, 2)
    kin = [0.13, -0.21, 0.7, 0.0, 0.0, 0.0]
    P.update_bodies(kin)
    xb, _ = P.body_state()
    S = _mk(small, 0, bodies=[(xb.copy(), np.asarray(small["bodies"][0].ds))])
    assert np.array_equal(P.body_idx(), S.body_idx())
    assert np.array_equal(P.weights(), S.weights())


def test_moving_off_the_grid_is_an_error(small):
    P = _mk(small, 2)
    with pytest.raises(ValueError):
        P.update_bodies([1.9, 0.0, 0.0, 0.0, 0.0, 0.0])


def test_cg_solves_the_schur_system(small, rng):
    P = _mk(small, 2)
    P.update_bodies([0.05, 0.03, 0.2, 0.0, 0.0, 0.0])
    nf = 2 * P.nbt
    B = np.column_stack([P.B_times(np.eye(nf)[:, j]) for j in range(nf)])
    b = rng.standard_normal(nf)
    x_ref = np.linalg.solve(B, b)
    x = P.cg(b)
    it, res, tol = P.cg_stats()
    assert 0 < it < 4 * nf  # B is near-symmetric (multigrid BCs), cond ~1e4: ~2 nf products
    assert res <= tol and tol == pytest.approx(1e-12 * np.linalg.norm(b), rel=1e-12)
    assert relerr(x, x_ref) < 1e-9
    # warm start at the solution: the first residual already meets the tolerance
    P.cg(b, x0=x_ref)
    assert P.cg_stats()[0] == 0


def test_body_at_rest_matches_static(small):
    """MotionFunction with xc = uc = 0 repeats the Static problem: E rebuilt
    every step at the same points, ub = 0, CG instead of the dense Binv."""
    Pm, Ps = _mk(small, 2), _mk(small, 0)
    Sm, Ss = Pm.new_state(_sc(Pm, U=1.0)), Ps.new_state(_sc(Ps, U=1.0))
    n = 12
    cdm, clm = Pm.run(Sm, np.tile(_sc(Pm, U=1.0, kin=np.zeros(6)), (n, 1)))
    cds, cls = Ps.run(Ss, np.tile(_sc(Ps, U=1.0), (n, 1)))
    assert relerr(cdm, cds) < 1e-9
    assert np.max(np.abs(clm - cls)) < 1e-9 * np.max(np.abs(cds))  # CL ~ 0: symmetric cylinder
    assert relerr(Sm.G, Ss.G) < 1e-9 and relerr(Sm.Fb, Ss.Fb) < 1e-9


def test_first_step_translation_equals_freestream(small):
    """From rest, a body moving at (-1, 0) in still fluid sees E(q) - ub h =
    h (1, 0), the right-hand side of a static body in a freestream Ux = 1
    (E of the uniform flux is U h up to Σδh = 1 +- 3.3e-9)."""
    Pm, Ps = _mk(small, 2), _mk(small, 0)
    Sm, Ss = Pm.new_state(_sc(Pm)), Ps.new_state(_sc(Ps, U=1.0))
    Pm.advance(Sm, _sc(Pm, kin=[0.0, 0.0, 0.0, -1.0, 0.0, 0.0]))
    Ps.advance(Ss, _sc(Ps, U=1.0))
    assert relerr(Sm.Fb, Ss.Fb) < 1e-7
    assert relerr(Sm.CD, Ss.CD) < 1e-7
    assert relerr(Sm.G, Ss.G) < 1e-7
    # a flipped sign or a missing h would be off by O(1)
    assert abs(Sm.CD[0]) > 1.0


def test_first_step_rotation_equals_moving_grid():
    """From rest, a plate spinning at ω about the origin (ub = ω (-y, x))
    against the MovingGrid frame spinning at θ̇ = ω (q0 = hc ω (Y, -X),
    timestepping.jl:284-298): E(q0) = -ub h up to the δh first-moment error."""
    c = case("small3")
    g = c["grid"]
    om = 0.8
    # a unit plate centred on the origin, points spaced h
    plate = [(b.xb, b.ds) for b in [ibpm.make_plate(1.0, 0.0, g.h, -0.5, 0.0, n=17)]]
    Pm = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], plate, motion=2)
    Pg = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], plate, motion=1)
    Sm, Sg = Pm.new_state(_sc(Pm)), Pg.new_state(_sc(Pg, mg=[0.0, 0.0, 0.0, om]))
    Pm.advance(Sm, _sc(Pm, kin=[0.0, 0.0, 0.0, 0.0, 0.0, om]))
    Pg.advance(Sg, _sc(Pg, mg=[0.0, 0.0, 0.0, om]))
    assert relerr(Sm.Fb, Sg.Fb) < 1e-6
    assert relerr(Sm.CL, Sg.CL) < 1e-6
    assert relerr(Sm.G[:, 0], Sg.G[:, 0]) < 1e-6
    assert np.linalg.norm(Sm.Fb) > 1e-3 * om * g.h
