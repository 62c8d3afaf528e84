"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (norm-relative, FP64): operators 1e-13 (same arithmetic up to the
FFT and summation order); Binv 1e-9 (the inverse amplifies the 1e-16
differences in B by cond(B) ~ 1e4-1e5); whole steps 1e-10, the bar of the
drop-in.  Index work (body_idx, E' accumulation order, rot!) is bit-exact.
"""
import math
import os

import numpy as np
import pytest

import ibpm
import oracle as O
from cases import boundary_of, case, gaussian_vortex_gamma, oracle_bodies, sc7
from conftest import relerr

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-10
OP_TOL = 1e-13


def build(name, share_binv=False):
    c = case(name)
    g = c["grid"]
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c),
                         motion=1 if "motion" in c else 0)
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"],
                        Binv=op.binv() if share_binv else None)
    return c, op, gp


@pytest.fixture(scope="module", params=["tiny", "small3", "mixed", "moving", "spec", "spec2", "direct", "single"])
def pair(request):
    return build(request.param)


def initial_states(c, op, gp, seed=7):
    """Same initial condition on both sides: IBState(prob), Γ = Gaussian
    vortex (+ seeded noise on level 1), then vort2flux! for q and ψ."""
    g = c["grid"]
    rng = np.random.default_rng(seed)
    G = gaussian_vortex_gamma(g, x0=0.8, y0=0.2, circ=0.5, width=0.3)
    G[:, 0] += 1e-3 * g.h * g.h * rng.standard_normal(g.nΓ)
    S = op.new_state(sc7(c, 0.0))
    psi, q, G2 = op.vort2flux(G)
    S.G[:] = G2
    S.psi[:] = psi
    S.q[:] = q
    st = ibpm.IBState(gp)
    st.Γ = G
    ibpm.vort2flux_(st, gp)
    return S, st


def compare_state(S, st, tol, fb_tol=None):
    """Norm-relative errors of the state fields; all must be <= tol (F̃b <=
    fb_tol when given)."""
    errs = {
        "q": relerr(st.q, S.q), "q0": relerr(st.q0, S.q0), "Γ": relerr(st.Γ, S.G), "ψ": relerr(st.ψ, S.psi),
        "nonlin1": relerr(st.nonlin[0], S.nl1), "nonlin2": relerr(st.nonlin[1], S.nl2),
        "F̃b": relerr(st.F̃b, S.Fb), "CD": relerr(st.CD, S.CD),
        # CL may vanish by symmetry: scale its error by the force magnitude
        "CL": np.max(np.abs(st.CL - S.CL)) / max(np.max(np.abs(S.CD)), np.max(np.abs(S.CL)), 1e-300),
    }
    bad = {k: v for k, v in errs.items() if not v <= (fb_tol if k == "F̃b" and fb_tol else tol)}
    assert not bad, f"fields beyond {tol}: {bad}"
    return errs


# ---------------------------------------------------------------------------
# setup products
# ---------------------------------------------------------------------------
def test_body_idx_bit_exact(pair):
    c, op, gp = pair
    assert np.array_equal(gp.body_idx, op.body_idx())


def test_binv(pair):
    c, op, gp = pair
    assert relerr(gp.Binv, op.binv()) < 1e-9
    assert gp.info.b_asymmetry < 1e-3


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
def _op(gp, name, *arrays, out_sizes):
    import ctypes
    from ibpm import _capi as C
    outs = [np.zeros(n) for n in out_sizes]
    args = [C.ptr(np.ascontiguousarray(a, dtype=np.float64)) for a in arrays] + [C.ptr(o) for o in outs]
    gp.ctx.check(getattr(C.lib(), name)(gp.h, *args), name)
    return outs


def test_dst2(pair, rng):
    c, op, gp = pair
    g = c["grid"]
    x = rng.standard_normal(g.nΓ)
    (y,) = _op(gp, "ibpm_op_dst2", x, out_sizes=[g.nΓ])
    assert relerr(y, op.dst2(x)) < OP_TOL


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_lapinv(pair, rng, kind):
    import ctypes
    from ibpm import _capi as C
    c, op, gp = pair
    g = c["grid"]
    for lev in range(1, g.mg + 1):
        b = rng.standard_normal(g.nΓ)
        x = np.zeros(g.nΓ)
        gp.ctx.check(C.lib().ibpm_op_lapinv(gp.h, kind, lev, C.ptr(b), C.ptr(x)))
        assert relerr(x, op.lapinv(b, kind, lev)) < OP_TOL, (kind, lev)


def test_nonlinear(pair, rng):
    from ibpm import _capi as C
    c, op, gp = pair
    g = c["grid"]
    q = np.asfortranarray(rng.standard_normal((g.nq, g.mg)))
    q0 = np.asfortranarray(rng.standard_normal((g.nq, g.mg)))
    G = np.asfortranarray(rng.standard_normal((g.nΓ, g.mg)))
    for lev in range(1, g.mg + 1):
        out = np.zeros(g.nΓ)
        gp.ctx.check(C.lib().ibpm_op_nonlinear(gp.h, lev, C.ptr(q), C.ptr(q0), C.ptr(G), C.ptr(out)))
        ref = op.nonlinear(q, q0, G, lev)
        assert np.array_equal(out, ref), (lev, relerr(out, ref))


def test_vort2flux(pair, rng):
    from ibpm import _capi as C
    c, op, gp = pair
    g = c["grid"]
    G = np.asfortranarray(rng.standard_normal((g.nΓ, g.mg)))
    Gd = G.copy(order="F")
    psi = np.zeros((g.nΓ, g.mg), order="F")
    q = np.zeros((g.nq, g.mg), order="F")
    gp.ctx.check(C.lib().ibpm_op_vort2flux(gp.h, C.ptr(Gd), C.ptr(psi), C.ptr(q)))
    psi_o, q_o, G_o = op.vort2flux(G)
    assert np.array_equal(Gd, G_o)  # coarsify! is exact arithmetic on the same inputs
    assert relerr(psi, psi_o) < 1e-12
    assert relerr(q, q_o) < 1e-12


def test_E_and_Et(pair, rng):
    c, op, gp = pair
    g = c["grid"]
    q = rng.standard_normal(g.nq)
    (fb,) = _op(gp, "ibpm_op_regT", q, out_sizes=[gp.nf])
    assert relerr(fb, op.regT(q)) < OP_TOL
    f = rng.standard_normal(gp.nf)
    (qq,) = _op(gp, "ibpm_op_reg", f, out_sizes=[g.nq])
    assert np.array_equal(qq, op.reg(f))  # same accumulation order as reg!


def test_rot_bit_exact(pair, rng):
    c, op, gp = pair
    g = c["grid"]
    q = rng.standard_normal(g.nq)
    (G,) = _op(gp, "ibpm_op_rot", q, out_sizes=[g.nΓ])
    assert np.array_equal(G, op.rot(q))


# ---------------------------------------------------------------------------
# whole steps
# ---------------------------------------------------------------------------
def test_state_init_matches(pair):
    c, op, gp = pair
    S = op.new_state(sc7(c, 0.0))
    st = ibpm.IBState(gp)
    assert np.array_equal(st.q0, S.q0)
    assert np.all(st.Γ == 0) and np.all(st.q == 0)


@pytest.mark.parametrize("name", ["tiny", "small3", "mixed", "moving"])
def test_one_step_shared_binv(name):
    c, op, gp = build(name, share_binv=True)
    S, st = initial_states(c, op, gp)
    t = 0.0
    op.advance(S, sc7(c, t))
    ibpm.advance_(st, gp, t)
    compare_state(S, st, 1e-12)


@pytest.mark.parametrize("name,nsteps", [("tiny", 40), ("small3", 30), ("mixed", 30), ("moving", 40), ("spec", 30),
                                          ("spec2", 20), ("direct", 30), ("single", 30)])
def test_many_steps(name, nsteps):
    c, op, gp = build(name)
    S, st = initial_states(c, op, gp)
    ts = np.arange(nsteps) * c["dt"]
    cd_o, cl_o = op.run(S, np.stack([sc7(c, t) for t in ts]))
    cd, cl, done = ibpm.advance_n_(st, gp, ts)
    assert done == nsteps
    assert relerr(cd, cd_o) < STEP_TOL
    # CL can sit near zero (symmetric flow): measure it against the force scale
    assert np.max(np.abs(cl - cl_o)) < STEP_TOL * np.abs(cd_o).max()
    compare_state(S, st, STEP_TOL)


@pytest.mark.parametrize("name", ["small3", "moving"])
def test_uploaded_q_is_the_state_q(name):
    """The state's q is what the next step's nonlinear! reads (timestepping.jl:
    96-98 through nonlinear.jl:168-201): a q uploaded that is not curl!(ψ) of
    the state's ψ must be used as uploaded.  One step from that state against
    the oracle, then further steps (whose q the FP64 step forms from ψ) and a
    download of q, all against the oracle."""
    c, op, gp = build(name, share_binv=True)
    S, st = initial_states(c, op, gp)
    rng = np.random.default_rng(11)
    dq = 1e-2 * np.max(np.abs(S.q)) * rng.standard_normal(S.q.shape)
    S.q[:] = S.q + dq
    st.q = st.q + dq
    assert relerr(st.q, S.q) < 1e-13
    ts = np.arange(4) * c["dt"]
    for k, t in enumerate(ts):
        op.advance(S, sc7(c, t))
        ibpm.advance_(st, gp, t)
        if k == 0:
            compare_state(S, st, 1e-12)
    compare_state(S, st, STEP_TOL)


def test_downloaded_q_equals_stored_curl():
    """After FP64 steps the library forms q when it is read; it must equal,
    bit for bit, curl! (curl.jl:81-116) of the state's ψ with the boundary
    values get_bc! (mg-utils.jl:7-32) of the next coarser level's, here
    through the oracle's curl! and get_bc!.  The step's ψ is vort2flux!(Γ)
    up to rounding (its chains read Ainv's transforms instead of
    transforming Γ again)."""
    c, op, gp = build("small3")
    g = c["grid"]
    _, st = initial_states(c, op, gp)
    ts = np.arange(6) * c["dt"]
    ibpm.advance_n_(st, gp, ts)
    q_formed = st.q.copy()
    psi = st.ψ.copy()
    for lev in range(g.mg):
        bc = op.get_bc(psi[:, lev + 1]) if lev + 1 < g.mg else np.zeros(op.nbc)
        assert np.array_equal(q_formed[:, lev], op.curl_mg(psi[:, lev], bc)), lev
    ibpm.vort2flux_(st, gp)  # the operator entry: Δinv chain and a curl! that stores q
    assert relerr(st.ψ, psi) < 1e-13


def test_advance_equals_advance_n():
    c, op, gp = build("tiny")
    _, st = initial_states(c, op, gp)
    ts = np.arange(5) * c["dt"]
    for t in ts:
        ibpm.advance_(st, gp, t)
    G1, q1, cd1 = st.Γ.copy(), st.q.copy(), st.CD.copy()
    gp2 = ibpm.IBProblem(c["grid"], c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=gp.Binv)
    _, st2 = initial_states(c, op, gp2)
    cd, cl, done = ibpm.advance_n_(st2, gp2, ts)
    assert np.array_equal(st2.Γ, G1) and np.array_equal(st2.q, q1) and cd[-1, 0] == cd1[0]


def test_in_step_profile_counts_and_leaves_results_unchanged():
    import ctypes
    from ibpm import _capi as C
    c, op, gp = build("tiny")
    _, st = initial_states(c, op, gp)
    ts = np.arange(6) * c["dt"]
    gp2 = ibpm.IBProblem(c["grid"], c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=gp.Binv)
    _, st2 = initial_states(c, op, gp2)
    gp2.ctx.check(C.lib().ibpm_set_option(gp2.h, C.OPT_PROFILE, 1))
    cd, _, _ = ibpm.advance_n_(st, gp, ts)
    cd2, _, _ = ibpm.advance_n_(st2, gp2, ts)
    ms, n = ctypes.c_double(0.0), ctypes.c_int64(0)
    gp2.ctx.check(C.lib().ibpm_profile_stats(gp2.h, ctypes.byref(ms), ctypes.byref(n)))
    assert n.value == len(ts) and ms.value > 0.0  # the Ainv y-pass of every step (A is a stencil)
    assert np.array_equal(cd, cd2) and np.array_equal(st.Γ, st2.Γ)


def test_restart_round_trip_is_bit_exact(tmp_path):
    """20 steps straight == 10 steps, restart file, fresh problem, 10 steps."""
    c, op, gp = build("small3", share_binv=True)
    _, st = initial_states(c, op, gp)
    ts = np.arange(20) * c["dt"]
    ibpm.advance_n_(st, gp, ts[:10])
    path = tmp_path / "restart.npz"
    ibpm.save_state(path, st, t=float(ts[9]))
    cd_a, cl_a, _ = ibpm.advance_n_(st, gp, ts[10:])
    gp2 = ibpm.IBProblem(c["grid"], c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=gp.Binv)
    st2 = ibpm.IBState(gp2)
    assert ibpm.load_state(path, st2) == float(ts[9])
    cd_b, cl_b, _ = ibpm.advance_n_(st2, gp2, ts[10:])
    assert np.array_equal(cd_a, cd_b) and np.array_equal(cl_a, cl_b)
    for name in ("Γ", "ψ", "q"):
        assert np.array_equal(getattr(st, name), getattr(st2, name)), name
    assert all(np.array_equal(a, b) for a, b in zip(st.nonlin, st2.nonlin))


def test_run_sim_saves_and_cfl():
    c, op, gp = build("tiny", share_binv=True)
    _, st = initial_states(c, op, gp)
    t = np.arange(30) * c["dt"]
    data = ibpm.init_save_data(t, {"save_fcns": [lambda tt, s: float(s.CD[0])], "save_times": [5 * c["dt"]]}, st)
    cd, cl = ibpm.run_sim_(t, st, gp, display_freq=25, data=data)
    assert cd.shape == (30, 1)
    assert np.allclose(data[0].t, t[::5])
    assert data[0].val == [float(v) for v in cd[::5, 0]]  # saved right after the step at each save time
    assert st.cfl > 0  # computed after step 25 (display_freq)


def test_cfl_matches_oracle_state():
    """compute_cfl = max|q[:,1]| dt / h (ibpm.jl:83-88), against the same
    maximum taken over the oracle's state after the same steps."""
    c, op, gp = build("tiny", share_binv=True)
    S, st = initial_states(c, op, gp)
    ts = np.arange(12) * c["dt"]
    op.run(S, np.stack([sc7(c, t) for t in ts]))
    ibpm.advance_n_(st, gp, ts)
    g = c["grid"]
    cfl_o = float(np.max(np.abs(S.q[:, 0]))) * c["dt"] / g.h
    cfl = ibpm.compute_cfl(st, gp)
    print(f"cfl {cfl!r} oracle {cfl_o!r}")
    assert cfl_o > 0.01 and abs(cfl - cfl_o) <= 1e-10 * cfl_o


def test_nonfinite_cl_stops_run_sim():
    """A NaN uploaded into Γ: the next step's CL is non-finite, ibpm_advance_n
    returns IBPM_E_NONFINITE after that step, run_sim! stops there
    (ibpm.jl:95-97) and compute_cfl reports NaN (the reference's maximum
    propagates it)."""
    import ctypes

    from ibpm import _capi as C

    c, op, gp = build("tiny", share_binv=True)
    _, st = initial_states(c, op, gp)
    G0 = np.array(st.Γ)
    G = G0.copy()
    G[G.shape[0] // 2, 0] = np.nan
    st.Γ = G
    t = np.arange(30) * c["dt"]
    arr, keep = gp.step_scalars(t[:5])
    cd = np.zeros((5, 1))
    cl = np.zeros((5, 1))
    done = ctypes.c_int(-1)
    rc = C.lib().ibpm_advance_n(gp.h, 5, C.ptr(t[:5]), arr, C.ptr(cd), C.ptr(cl), ctypes.byref(done))
    st.invalidate()
    assert rc == C.IBPM_E_NONFINITE and done.value == 1 and not np.isfinite(cl[0, 0])
    assert math.isnan(ibpm.compute_cfl(st, gp))
    st.Γ = G
    ibpm.vort2flux_(st, gp)
    cd, cl = ibpm.run_sim_(t, st, gp, display_freq=25)
    assert cl.shape == (1, 1) and not np.isfinite(cl[0, 0])
    # the AB2 history now holds NaN too; a fresh state steps on the same problem
    st = ibpm.IBState(gp)
    st.Γ = G0
    ibpm.vort2flux_(st, gp)
    cd, cl = ibpm.run_sim_(t, st, gp, display_freq=25)
    assert cl.shape == (30, 1) and np.all(np.isfinite(cl))


@pytest.mark.parametrize("name", ["small3", "tiny"])
def test_autotune_applies_legal_shapes(name):
    """The autotuner sweeps its options on a problem (one-level problems
    time the y-pass without boundary terms) and leaves each at a value it
    timed; stepping afterwards still works."""
    import ctypes
    from ibpm import _capi as C
    from ibpm.autotune import SWEEPS, autotune
    c = case(name)
    gp = ibpm.IBProblem(c["grid"], c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    res = autotune(gp, reps=2)
    assert set(res) == set(SWEEPS)
    for opt, (best, times) in res.items():
        assert best in times and all(t > 0.0 for t in times.values())
        v = ctypes.c_int(-1)
        gp.ctx.check(C.lib().ibpm_get_option(gp.h, opt, ctypes.byref(v)))
        assert v.value == best
    st = ibpm.IBState(gp)
    _, _, done = ibpm.advance_n_(st, gp, np.arange(2) * c["dt"])
    assert done == 2


@pytest.mark.parametrize("name", ["small3", "spec", "cyl_re40_200x200x3"])
def test_launch_shapes_give_identical_steps(name):
    """Every legal work-group shape of the x and y passes (the autotuner's
    options) transforms each line with the same arithmetic: the steps are
    bit-identical to the default shape's."""
    from ibpm import _capi as C
    c = case(name)
    g = c["grid"]
    ts = np.arange(3) * c["dt"]

    def run(opts):
        gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=ref_binv)
        for o, v in opts.items():
            if C.lib().ibpm_set_option(gp.h, o, v) != C.IBPM_OK:
                return None  # does not fit this grid
        st = ibpm.IBState(gp)
        st.Γ = gaussian_vortex_gamma(g, x0=0.8, y0=0.2, circ=0.5, width=0.3)
        ibpm.vort2flux_(st, gp)
        cd, cl, _ = ibpm.advance_n_(st, gp, ts)
        return cd, st.Γ.copy(), st.q.copy()

    ref_binv = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"]).Binv
    base = run({})
    n = 0
    for yb in (4, 8, 12):
        for ys in (4, 8, 12):
            for xw in (4, 8, 12, 14):
                r = run({C.OPT_YTILE_BATCH: yb, C.OPT_YTILE_SINGLE: ys, C.OPT_XWAVES_BATCH: xw})
                if r is None:
                    continue
                n += 1
                for a, b in zip(r, base):
                    assert np.array_equal(a, b), (yb, ys, xw)
    assert n >= 12


def test_graph_replays_match_direct_launches():
    """Captured steps (default) against a twin problem that launches every
    kernel: several rounds of replays with eager work and syncs in between,
    every round compared bit for bit."""
    from ibpm import _capi as C
    pairs = [build("small3", share_binv=True) for _ in range(2)]
    gps = [gp for _, _, gp in pairs]
    gps[1].ctx.check(C.lib().ibpm_set_option(gps[1].h, C.OPT_GRAPHS, 0))
    sts = [initial_states(c, op, gp)[1] for c, op, gp in pairs]
    c = pairs[0][0]
    g = c["grid"]
    ts = np.arange(31) * c["dt"]
    for k in range(3):
        cds = [ibpm.advance_n_(st, gp, ts[10 * k: 10 * k + 10])[0] for st, gp in zip(sts, gps)]
        assert np.array_equal(cds[0], cds[1]), k
        for name in ("Γ", "q", "ψ"):
            assert np.array_equal(getattr(sts[0], name), getattr(sts[1], name)), (k, name)
        # eager work on the same streams between the replays
        for st, gp in zip(sts, gps):
            ibpm.compute_cfl(st, gp)
            _op(gp, "ibpm_op_dst2", st.Γ[:, 0], out_sizes=[g.nΓ])
    for st, gp in zip(sts, gps):
        ibpm.advance_(st, gp, ts[-1])
    assert np.array_equal(sts[0].Γ, sts[1].Γ) and np.array_equal(sts[0].CD, sts[1].CD)


def test_batch_of_problems_equals_one_by_one():
    """Concurrent streams (ibpm_advance_async / _wait) change nothing."""
    names = ["tiny", "mixed", "small3"]
    runs = []
    for batched in (False, True):
        pairs = [build(n, share_binv=True) for n in names]
        states = [initial_states(c, op, gp)[1] for c, op, gp in pairs]
        tss = [np.arange(12) * c["dt"] for c, _, _ in pairs]
        if batched:
            res = ibpm.advance_batch_(states, [gp for _, _, gp in pairs], tss)
        else:
            res = [ibpm.advance_n_(st, gp, ts) for st, (_, _, gp), ts in zip(states, pairs, tss)]
        runs.append((res, [st.Γ.copy() for st in states]))
    (ra, ga), (rb, gb) = runs
    for (cda, cla, na), (cdb, clb, nb) in zip(ra, rb):
        assert na == nb == 12 and np.array_equal(cda, cdb) and np.array_equal(cla, clb)
    assert all(np.array_equal(a, b) for a, b in zip(ga, gb))


# ---------------------------------------------------------------------------
# LinearizedIBProblem (problem-types.jl:47-87)
# ---------------------------------------------------------------------------
def _linearized(name, nbase=6):
    """GPU base state after a few steps, the linearized problem on both sides."""
    c, op, gp = build(name, share_binv=True)
    _, st = initial_states(c, op, gp)
    ibpm.advance_n_(st, gp, np.arange(nbase) * c["dt"])
    g = c["grid"]
    lp = ibpm.LinearizedIBProblem(st, gp, c["dt"])
    olp = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=gp.Binv)
    olp.linearize(st.q, st.q0, st.Γ)
    return c, lp, olp


@pytest.mark.parametrize("name", ["tiny", "small3", "spec"])
def test_linearized_nonlinear_bit_exact(name, rng):
    from ibpm import _capi as C
    c, lp, olp = _linearized(name)
    g = c["grid"]
    q = np.asfortranarray(rng.standard_normal((g.nq, g.mg)))
    q0 = np.zeros((g.nq, g.mg), order="F")
    G = np.asfortranarray(rng.standard_normal((g.nΓ, g.mg)))
    for lev in range(1, g.mg + 1):
        out = np.zeros(g.nΓ)
        lp.ctx.check(C.lib().ibpm_op_nonlinear(lp.h, lev, C.ptr(q), C.ptr(q0), C.ptr(G), C.ptr(out)))
        ref = olp.nonlinear(q, q0, G, lev)
        assert np.array_equal(out, ref), (lev, relerr(out, ref))


def test_linearized_steps():
    c, lp, olp = _linearized("small3")
    g = c["grid"]
    st = ibpm.IBState(lp)
    assert np.all(st.q0 == 0.0)  # base_flux! of the linearized problem
    Gp = 1e-3 * gaussian_vortex_gamma(g, x0=0.9, y0=0.3, circ=0.5, width=0.3)
    S = olp.new_state(np.zeros(7))
    psi, q, G2 = olp.vort2flux(Gp)
    S.G[:] = G2
    S.psi[:] = psi
    S.q[:] = q
    st.Γ = Gp
    ibpm.vort2flux_(st, lp)
    ts = np.arange(20) * c["dt"]
    cd_o, cl_o = olp.run(S, np.zeros((len(ts), 7)))
    cd, cl, done = ibpm.advance_n_(st, lp, ts)
    assert done == len(ts)
    assert relerr(cd, cd_o) < STEP_TOL
    compare_state(S, st, STEP_TOL)


def test_cyl_re40_200x200_steps():
    """The Re 40 cylinder from an impulsive start (200 x 200, 3 levels), 20 steps."""
    c, op, gp = build("cyl_re40_200x200x3")
    assert relerr(gp.Binv, op.binv()) < 1e-9
    S = op.new_state(sc7(c, 0.0))
    st = ibpm.IBState(gp)
    ts = np.arange(20) * c["dt"]
    cd_o, cl_o = op.run(S, np.stack([sc7(c, t) for t in ts]))
    cd, cl, _ = ibpm.advance_n_(st, gp, ts)
    assert relerr(cd, cd_o) < STEP_TOL
    assert np.max(np.abs(cl - cl_o)) < 1e-10 * np.abs(cd_o).max()
    compare_state(S, st, STEP_TOL)


# ---------------------------------------------------------------------------
# full-size properties (1024 x 1024, 5 levels), no oracle run needed
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def big():
    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    nf = 2 * len(c["bodies"][0].ds)
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=np.eye(nf))
    return c, gp


def test_full_size_dst_involution(big, rng):
    c, gp = big
    g = c["grid"]
    x = rng.standard_normal(g.nΓ)
    (y,) = _op(gp, "ibpm_op_dst2", x, out_sizes=[g.nΓ])
    (z,) = _op(gp, "ibpm_op_dst2", y, out_sizes=[g.nΓ])
    assert relerr(z / (4.0 * g.nx * g.ny), x) < 1e-13


def test_full_size_dinv_inverts_5pt(big, rng):
    from ibpm import _capi as C
    c, gp = big
    g = c["grid"]
    n1, n2 = g.nx - 1, g.ny - 1
    psi = rng.standard_normal((n1, n2))
    pad = np.zeros((n1 + 2, n2 + 2))
    pad[1:-1, 1:-1] = psi
    lap = 4 * psi - pad[2:, 1:-1] - pad[:-2, 1:-1] - pad[1:-1, 2:] - pad[1:-1, :-2]
    x = np.zeros(g.nΓ)
    b = np.ascontiguousarray(lap.ravel(order="F"))
    gp.ctx.check(C.lib().ibpm_op_lapinv(gp.h, 0, 1, C.ptr(b), C.ptr(x)))
    assert relerr(x, psi.ravel(order="F")) < 1e-11


def test_full_size_b_columns_and_binv():
    """Columns of the Schur matrix B at 1024 x 1024 x 5 (804 unknowns):
    B_times! (laplacian.jl:143-169) of 8 unit forces through the GPU's
    operators (E', rot!, Ainv, vort2flux!, E) against the oracle's B_times!,
    to 1e-12; and the GPU's Binv (B assembled on the GPU, inverted on the
    host) applied to the oracle's columns gives the unit vectors back, to the
    level cond(B) ~ 5e5 allows."""
    import ctypes

    from ibpm import _capi as C

    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    nf = gp.nf
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=np.eye(nf))
    O.set_threads(16)
    Binv = gp.Binv
    L = C.lib()
    worst_b, worst_r = 0.0, 0.0
    for j in (0, 1, nf // 4, nf // 2 - 1, nf // 2, nf // 2 + 7, 3 * nf // 4, nf - 1):
        e = np.zeros(nf)
        e[j] = 1.0
        b_o = op.B_times(e)
        (q,) = _op(gp, "ibpm_op_reg", e, out_sizes=[g.nq])
        (G1,) = _op(gp, "ibpm_op_rot", q, out_sizes=[g.nΓ])
        x = np.zeros(g.nΓ)
        gp.ctx.check(L.ibpm_op_lapinv(gp.h, 2, 1, C.ptr(G1), C.ptr(x)))
        G = np.zeros((g.nΓ, g.mg), order="F")
        G[:, 0] = x
        psi = np.zeros((g.nΓ, g.mg), order="F")
        qa = np.zeros((g.nq, g.mg), order="F")
        gp.ctx.check(L.ibpm_op_vort2flux(gp.h, C.ptr(G), C.ptr(psi), C.ptr(qa)))
        (b,) = _op(gp, "ibpm_op_regT", qa[:, 0].copy(), out_sizes=[nf])
        r = Binv @ b_o
        r[j] -= 1.0
        worst_b = max(worst_b, relerr(b, b_o))
        worst_r = max(worst_r, float(np.linalg.norm(r)))
    print(f"B columns: worst relative difference {worst_b:.2e}; |Binv B e_j - e_j| worst {worst_r:.2e}")
    assert worst_b < 1e-12
    assert worst_r < 1e-7


# ---------------------------------------------------------------------------
# full-size steps against the oracle: the 1024 x 1024 and 800 x 400 grids,
# 5 levels, and the 512 x 256, 4-level moving grid (the 1023-point transform
# core, the compile-time cores and the vort2flux! boundary terms at these widths)
# ---------------------------------------------------------------------------
def _smooth_along_body(f):
    """(1/4, 1/2, 1/4) filter along the body points of each force component
    (F̃b holds all x components, then all y components)."""
    out = []
    for comp in np.split(np.asarray(f, dtype=np.float64).ravel(), 2):
        out.append(0.25 * np.roll(comp, 1) + 0.5 * comp + 0.25 * np.roll(comp, -1))
    return np.concatenate(out)


@pytest.mark.parametrize("name,nsteps", [("cyl_re200_1024x1024x5", 10), ("cyl_re100_800x400x5", 10), ("plate_pitch_512x256x4", 10)])
def test_full_size_steps_match_oracle(name, nsteps):
    """vort2flux! and advance! steps at full size.  The oracle takes the
    GPU's Binv: assembling B column by column would take the CPU minutes here,
    and Binv itself is compared with the oracle's at the smaller sizes.

    F̃b = Binv (E q - u_b) pointwise is ill-conditioned at these sizes
    (cond(B) = 5.0e5 at 1024 x 1024): the 2e-12 rounding difference of q
    between two FP64 transform orders reaches F̃b as up to 6e-10 relative
    over 40 steps, all of it in modes that alternate along the body.  The
    quantities built from the forces (CD, CL) and the forces smoothed along
    the body are held to the 1e-10 bar; raw F̃b to 1e-8."""
    c = case(name)
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c),
                         motion=1 if "motion" in c else 0, binv=gp.Binv)
    O.set_threads(min(8, os.cpu_count() or 1))
    try:
        S, st = initial_states(c, op, gp)
        assert np.array_equal(st.Γ, S.G) and relerr(st.ψ, S.psi) < 1e-12
        # q is differences of neighbouring ψ: at h = 1/128 they are ~1e-3 of ψ,
        # so ψ's rounding shows in q about that much larger (8e-12 measured)
        assert relerr(st.q, S.q) < STEP_TOL
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
    print(name, {k: f"{v:.1e}" for k, v in errs.items()})
    assert errs["F̃b smoothed"] < STEP_TOL


@pytest.mark.skipif(not os.environ.get("IBPM_LONG_STEPS"),
                    reason="long full-size run: set IBPM_LONG_STEPS (e.g. 300) to run it")
def test_full_size_long_run():
    """IBPM_LONG_STEPS steps of the 1024 x 1024, 5-level cylinder against the
    oracle (GPU's Binv), compared every 25 steps: CD, CL, Γ, ψ and the forces
    smoothed along the body to 1e-10, raw F̃b to 1e-8."""
    nsteps = int(os.environ["IBPM_LONG_STEPS"])
    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=gp.Binv)
    O.set_threads(min(15, os.cpu_count() or 1))
    try:
        S, st = initial_states(c, op, gp)
        k = 0
        while k < nsteps:
            ts = (k + np.arange(25)) * c["dt"]
            cd_o, cl_o = op.run(S, np.stack([sc7(c, t) for t in ts]))
            cd, cl, _ = ibpm.advance_n_(st, gp, ts)
            k += 25
            errs = compare_state(S, st, STEP_TOL, fb_tol=1e-8)
            errs["CD run"] = relerr(cd, cd_o)
            errs["CL run"] = np.max(np.abs(cl - cl_o)) / np.abs(cd_o).max()
            errs["F̃b smoothed"] = relerr(_smooth_along_body(st.F̃b), _smooth_along_body(S.Fb))
            print(k, {n: f"{v:.1e}" for n, v in errs.items()}, flush=True)
            assert errs["CD run"] < STEP_TOL and errs["CL run"] < STEP_TOL and errs["F̃b smoothed"] < STEP_TOL
    finally:
        O.set_threads(1)
