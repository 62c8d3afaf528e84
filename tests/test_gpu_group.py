"""Groups (ibpm_group_*): problems on one grid stepped together, every pass of
advance! one launch over all of them.  Each member's arithmetic is the one of
stepping it alone, so the results must be identical bit for bit: forces,
every state field, and the continuation after the group is dissolved.
"""
import numpy as np
import pytest

import ibpm
from cases import gaussian_vortex_gamma, naca64_cases
from ibpm import _capi as C

pytestmark = pytest.mark.gpu


def members(binvs=None):
    """Three problems on the 64 x 64, 3-level grid with different bodies (one
    with two bodies), Reynolds numbers, time steps and freestreams."""
    g = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -2.0, 2.0), mg=3)
    specs = [
        ([ibpm.make_cylinder(0.5, g.h, 0.0, 0.0, n=40)], 100.0, 0.01, {"Ux": lambda t: 1.0}),
        ([ibpm.make_plate(1.0, 0.3, g.h, -0.4, 0.1, n=17)], 200.0, 0.005,
         {"Ux": lambda t: 1.0, "inclination": lambda t: 0.2}),
        ([ibpm.make_cylinder(0.25, g.h, -0.3, 0.2, n=20), ibpm.make_plate(0.5, 0.0, g.h, 0.4, -0.3, n=9)], 50.0,
         0.008, {"Ux": lambda t: 0.8, "Uy": lambda t: 0.1}),
    ]
    probs, states = [], []
    for k, (bodies, Re, dt, fs) in enumerate(specs):
        p = ibpm.IBProblem(g, bodies, dt, Re, freestream=fs, Binv=None if binvs is None else binvs[k])
        st = ibpm.IBState(p)
        st.Γ = gaussian_vortex_gamma(g, x0=0.8, y0=0.2 * k, circ=0.5, width=0.3)
        ibpm.vort2flux_(st, p)
        probs.append(p)
        states.append(st)
    return probs, states


def fields(st):
    return [st.Γ.copy(), st.q.copy(), st.ψ.copy(), st.F̃b.copy(), st.nonlin[0].copy(), st.nonlin[1].copy()]


def test_group_steps_equal_single_steps():
    pa, sa = members()
    pg, sg = members([p.Binv for p in pa])
    n1, n2 = 12, 5
    alone = [ibpm.advance_n_(st, p, np.arange(n1) * p.scheme.dt) for p, st in zip(pa, sa)]
    grp = ibpm.Group(pg, sg)
    res = grp.advance_n_(n1)
    for (cd, cl, d), (cdg, clg, dg), st, stg in zip(alone, res, sa, sg):
        assert d == dg == n1
        assert np.array_equal(cd, cdg) and np.array_equal(cl, clg)
        for a, b in zip(fields(st), fields(stg)):
            assert np.array_equal(a, b)
    # members step through the group only
    with pytest.raises(C.IBPMError):
        ibpm.advance_(sg[0], pg[0], 0.0)
    # dissolved: every member carries on alone from the same state
    grp.close()
    for p, st, pgi, sgi in zip(pa, sa, pg, sg):
        cd, cl, _ = ibpm.advance_n_(st, p, np.arange(n2) * p.scheme.dt)
        cdg, clg, _ = ibpm.advance_n_(sgi, pgi, np.arange(n2) * pgi.scheme.dt)
        assert np.array_equal(cd, cdg) and np.array_equal(cl, clg)
        assert np.array_equal(st.Γ, sgi.Γ) and np.array_equal(st.q, sgi.q)


def test_group_of_naca_instances_equals_single():
    """Instances on the same finest grid with 3, 4 and 5 levels in one group:
    the group keeps every instance at 5 levels and skips the levels an
    instance does not have."""
    allc = naca64_cases()
    nx0 = allc[0]["grid"].nx
    same = [c for c in allc if c["grid"].nx == nx0]
    cs = []
    for mg in (3, 4, 5):
        cs += [c for c in same if c["grid"].mg == mg][:2]
    assert sorted({c["grid"].mg for c in cs}) == [3, 4, 5]

    def build():
        out = []
        for c in cs:
            p = ibpm.IBProblem(c["grid"], c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
            out.append((p, ibpm.IBState(p)))
        return out

    a, b = build(), build()
    alone = [ibpm.advance_n_(st, p, np.arange(6) * p.scheme.dt) for p, st in a]
    grp = ibpm.Group([p for p, _ in b], [st for _, st in b])
    res = grp.advance_n_(3)
    res2 = grp.advance_n_(3)  # two calls: the history slots restart at each wait
    for (cd, cl, _), (c1, l1, _), (c2, l2, _), (_, sta), (_, stb) in zip(alone, res, res2, a, b):
        assert np.array_equal(cd, np.vstack([c1, c2])) and np.array_equal(cl, np.vstack([l1, l2]))
        assert np.array_equal(sta.Γ, stb.Γ)
    grp.close()


def test_group_with_fewer_levels_member():
    """A 2-level member beside 3-level ones (same finest grid)."""
    pa, sa = members()
    pg, sg = members([p.Binv for p in pa])
    g2 = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -2.0, 2.0), mg=2)

    def two_level():
        p = ibpm.IBProblem(g2, [ibpm.make_cylinder(0.4, g2.h, 0.1, 0.0, n=30)], 0.01, 80.0,
                           freestream={"Ux": lambda t: 1.0})
        st = ibpm.IBState(p)
        st.Γ = gaussian_vortex_gamma(g2, x0=0.7, y0=0.1, circ=0.5, width=0.3)
        ibpm.vort2flux_(st, p)
        return p, st

    p2a, s2a = two_level()
    p2g, s2g = two_level()
    alone = [ibpm.advance_n_(st, p, np.arange(8) * p.scheme.dt) for p, st in zip(pa + [p2a], sa + [s2a])]
    grp = ibpm.Group(pg + [p2g], sg + [s2g])
    res = grp.advance_n_(8)
    for (cd, cl, _), (cdg, clg, _), st, stg in zip(alone, res, sa + [s2a], sg + [s2g]):
        assert np.array_equal(cd, cdg) and np.array_equal(cl, clg)
        for x, y in zip(fields(st), fields(stg)):
            assert np.array_equal(x, y)
    grp.close()


def test_group_rejects_mixed_grids():
    pa, _ = members()
    g2 = ibpm.make_grid(1.0 / 16, (-1.5, 2.5, -1.0, 1.0), mg=3)
    other = ibpm.IBProblem(g2, [ibpm.make_cylinder(0.3, g2.h, 0.0, 0.0, n=24)], 0.01, 100.0)
    with pytest.raises(C.IBPMError):
        ibpm.Group([pa[0], other])


def test_group_graph_replay_equals_launches():
    """A group's steps replayed from captured graphs (the default) are the
    same launches as stepping it without graphs: bit-identical."""
    pa, sa = members()
    pg, sg = members([p.Binv for p in pa])
    ga = ibpm.Group(pa, sa)
    ga.set_option(C.OPT_GRAPHS, 0)
    gb = ibpm.Group(pg, sg)
    ra = ga.advance_n_(9)
    rb = gb.advance_n_(9)
    rb2 = gb.advance_n_(4)  # graphs replayed with the roles of later steps
    ra2 = ga.advance_n_(4)
    for (cd, cl, d), (cdg, clg, dg) in zip(ra + ra2, rb + rb2):
        assert d == dg and np.array_equal(cd, cdg) and np.array_equal(cl, clg)
    ga.close()
    gb.close()
    for st, stg in zip(sa, sg):
        for a, b in zip(fields(st), fields(stg)):
            assert np.array_equal(a, b)


def test_joint_graph_of_three_naca_shapes_matches_oracle():
    """One NACA 0012 instance of each grid shape of the batched workload
    (400 x 200 with 3 levels, 600 x 300 with 4, 800 x 400 with 5), each in
    its own group, the three groups stepped by one graph per step for 20
    steps from the impulsive start: every instance against the oracle (which
    takes the GPU's Binv) to the 1e-10 bar."""
    import oracle as O
    from cases import boundary_of, oracle_bodies, sc7
    from conftest import relerr

    want = [(400, 200, 3), (600, 300, 4), (800, 400, 5)]
    cs = naca64_cases()
    picked = [next(c for c in cs if (c["grid"].nx, c["grid"].ny, c["grid"].mg) == w) for w in want]
    O.set_threads(16)
    n = 20
    probs, states, oracles = [], [], []
    for c in picked:
        g = c["grid"]
        p = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
        st = ibpm.IBState(p)
        op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=p.Binv)
        S = op.new_state(sc7(c, 0.0))
        cd_o, cl_o = op.run(S, np.stack([sc7(c, k * c["dt"]) for k in range(n)]))
        probs.append(p)
        states.append(st)
        oracles.append((S, cd_o, cl_o))
    groups = [ibpm.Group([p], [st]) for p, st in zip(probs, states)]
    res = ibpm.advance_groups_(groups, n)
    for grp in groups:
        grp.close()
    for c, (r,), st, (S, cd_o, cl_o) in zip(picked, res, states, oracles):
        cd, cl, d = r
        assert d == n
        scale = max(np.max(np.abs(cd_o)), np.max(np.abs(cl_o)))
        errs = {"CD": float(np.max(np.abs(cd - cd_o)) / scale), "CL": float(np.max(np.abs(cl - cl_o)) / scale),
                "Γ": relerr(st.Γ, S.G), "ψ": relerr(st.ψ, S.psi), "q": relerr(st.q, S.q)}
        g = c["grid"]
        print(f"{g.nx}x{g.ny}x{g.mg}: final CD {cd[-1, 0]:.8f} CL {cl[-1, 0]:.3e};", errs)
        assert all(v <= 1e-10 for v in errs.values()), errs


def test_autotune_group_keeps_results():
    """autotune_group picks the all-level launch shapes of a group by timing
    group steps; the shapes change no result: the tuned group's state equals
    an untuned group's after the same number of steps, bit for bit."""
    from ibpm.autotune import autotune_group
    pa, sa = members()
    pg, sg = members([p.Binv for p in pa])
    ga, gb = ibpm.Group(pa, sa), ibpm.Group(pg, sg)
    taken = []
    orig = gb.advance_async
    gb.advance_async = lambda n: (taken.append(n), orig(n))[1]
    chosen = autotune_group(gb, steps=2)
    assert set(chosen) == {C.OPT_YTILE_BATCH, C.OPT_PIPE, C.OPT_XWAVES_BATCH}
    for o, v in chosen.items():
        assert pg[0].get_option(o) == v
    ga.advance_async(sum(taken))
    ga.wait()
    ga.advance_async(3)
    gb.advance_async(3)
    for (cda, cla, da), (cdb, clb, db) in zip(ga.wait(), gb.wait()):
        assert da == db == 3
        assert np.array_equal(cda, cdb) and np.array_equal(cla, clb)
    ga.close()
    gb.close()
    for st, stg in zip(sa, sg):
        for a, b in zip(fields(st), fields(stg)):
            assert np.array_equal(a, b)
