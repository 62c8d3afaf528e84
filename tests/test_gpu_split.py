"""The faster forms of the step against the oracle.

The one-level y-pass of 1023-point columns in both forms (IBPM_OPT_YSPLIT):
1 (the default), each column transformed by two wavefronts at once, the even
outputs as a 511-point DST-I of the differences x_j - x_{M-j} and the odd
outputs as a DCT-III of the sums; 0, one wavefront per column.  Against the
oracle with the tolerances of test_gpu_parity's full-size cases: operators
1e-11/1e-12, steps 1e-10.  And the vort2flux! chains that read Γ's
x-transform back from Ainv's output (IBPM_OPT_XREUSE) on a small grid, where
it is off by default, alone and in a group."""
import os

import numpy as np
import pytest

import ibpm
import oracle as O
from cases import boundary_of, case, oracle_bodies, sc7
from conftest import relerr
from test_gpu_parity import STEP_TOL, _smooth_along_body, compare_state, initial_states

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    nf = 2 * len(c["bodies"][0].ds)
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=np.eye(nf))
    return c, gp


@pytest.mark.parametrize("ysplit", [1, 0])
def test_split_dinv_inverts_5pt(big, rng, ysplit):
    """Δinv of one level (a single-level launch)
    inverts the 5-point Laplacian with zero boundary values: the transform
    pair, the eigenvalue scaling and both halves of each transform."""
    from ibpm import _capi as C
    c, gp = big
    g = c["grid"]
    gp.set_option(C.OPT_YSPLIT, ysplit)
    try:
        n1, n2 = g.nx - 1, g.ny - 1
        psi = rng.standard_normal((n1, n2))
        pad = np.zeros((n1 + 2, n2 + 2))
        pad[1:-1, 1:-1] = psi
        lap = 4 * psi - pad[2:, 1:-1] - pad[:-2, 1:-1] - pad[1:-1, 2:] - pad[1:-1, :-2]
        b = np.ascontiguousarray(lap.ravel(order="F"))
        x = np.zeros(g.nΓ)
        gp.ctx.check(C.lib().ibpm_op_lapinv(gp.h, 0, 1, C.ptr(b), C.ptr(x)))
        assert relerr(x, psi.ravel(order="F")) < 1e-11
    finally:
        gp.set_option(C.OPT_YSPLIT, 1)


@pytest.mark.parametrize("ysplit", [1, 0])
def test_split_vort2flux_matches_oracle(big, rng, ysplit):
    """vort2flux! of a random Γ at 1024 x 1024 x 5: every level's y-pass with
    the boundary terms of the next coarser ψ runs split."""
    from ibpm import _capi as C
    c, gp = big
    g = c["grid"]
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=np.eye(gp.nf))
    G = np.asfortranarray(rng.standard_normal((g.nΓ, g.mg)))
    O.set_threads(min(8, os.cpu_count() or 1))
    try:
        psi_o, q_o, G_o = op.vort2flux(G)
    finally:
        O.set_threads(1)
    gp.set_option(C.OPT_YSPLIT, ysplit)
    try:
        Gd = G.copy(order="F")
        psi = np.zeros((g.nΓ, g.mg), order="F")
        q = np.zeros((g.nq, g.mg), order="F")
        gp.ctx.check(C.lib().ibpm_op_vort2flux(gp.h, C.ptr(Gd), C.ptr(psi), C.ptr(q)))
    finally:
        gp.set_option(C.OPT_YSPLIT, 1)
    assert np.array_equal(Gd, G_o)
    assert relerr(psi, psi_o) < 1e-12
    assert relerr(q, q_o) < 1e-11


@pytest.mark.parametrize("ysplit", [1, 0])
def test_split_steps_match_oracle(ysplit):
    """10 advance! steps of the 1024 x 1024, 5-level cylinder (graph replay
    on) against the oracle, with either form of the one-level y-passes."""
    from ibpm import _capi as C
    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    gp.set_option(C.OPT_YSPLIT, ysplit)
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=gp.Binv)
    O.set_threads(min(8, os.cpu_count() or 1))
    nsteps = 10
    try:
        S, st = initial_states(c, op, gp)
        ts = np.arange(nsteps) * c["dt"]
        cd_o, cl_o = op.run(S, np.stack([sc7(c, t) for t in ts]))
    finally:
        O.set_threads(1)
    cd, cl, done = ibpm.advance_n_(st, gp, ts)
    assert done == nsteps
    assert relerr(cd, cd_o) < STEP_TOL
    assert np.max(np.abs(cl - cl_o)) < STEP_TOL * np.abs(cd_o).max()
    errs = compare_state(S, st, STEP_TOL, fb_tol=1e-8)
    errs["F̃b smoothed"] = relerr(_smooth_along_body(st.F̃b), _smooth_along_body(S.Fb))
    print("ysplit", ysplit, {k: f"{v:.1e}" for k, v in errs.items()})
    assert errs["F̃b smoothed"] < STEP_TOL


@pytest.mark.parametrize("name,xreuse", [("small3", 1), ("moving", 1), ("small3", 0)])
def test_xreuse_steps_match_oracle(name, xreuse):
    """IBPM_OPT_XREUSE on a small grid (where it is off by default): the
    vort2flux! chains of the step x-transform only the lines coarsify!
    rewrites and read the rest from Ainv's y-pass output; 30 steps against
    the oracle, static and moving grid."""
    from ibpm import _capi as C
    from test_gpu_parity import build
    c, op, gp = build(name, share_binv=True)
    gp.set_option(C.OPT_XREUSE, xreuse)
    S, st = initial_states(c, op, gp)
    ts = np.arange(30) * c["dt"]
    cd_o, cl_o = op.run(S, np.stack([sc7(c, t) for t in ts]))
    cd, cl, done = ibpm.advance_n_(st, gp, ts)
    assert done == 30
    assert relerr(cd, cd_o) < STEP_TOL
    compare_state(S, st, STEP_TOL)


def test_group_equals_single_with_xreuse():
    """Groups follow their problems' IBPM_OPT_XREUSE and stay bit-identical to
    the members stepped alone with it on."""
    from ibpm import _capi as C
    from test_gpu_group import fields, members
    pa, sa = members()
    pg, sg = members([p.Binv for p in pa])
    for p in pa + pg:
        p.set_option(C.OPT_XREUSE, 1)
    n1 = 8
    alone = [ibpm.advance_n_(st, p, np.arange(n1) * p.scheme.dt) for p, st in zip(pa, sa)]
    grp = ibpm.Group(pg, sg)
    res = grp.advance_n_(n1)
    for (cd, cl, d), (cdg, clg, dg), st, stg in zip(alone, res, sa, sg):
        assert d == dg == n1
        assert np.array_equal(cd, cdg) and np.array_equal(cl, clg)
        for a, b in zip(fields(st), fields(stg)):
            assert np.array_equal(a, b)
    grp.close()
