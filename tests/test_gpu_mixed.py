"""The mixed-precision form (IBPM_OPT_PRECISION) against the FP64 oracle.

Every DST-I direction of Δinv and Ainv runs as a dense sine-matrix product on
the MFMA units, in FP32 (1) or split BF16 (2: high and low parts, three
products, FP32 accumulation), on FP64 residuals of the state; the eigenvalue
scaling, residuals, updates, the Schur solve and the force sums stay FP64.

Bars (the benchmark's "mixed @ CD err <= 1e-6"): CD and CL within 1e-6 of the
FP64 oracle's, relative to the largest |CD| of the run, at every step.  The
errors of ω (Γ / hc²), ψ and q are printed for the record.
"""
import numpy as np
import pytest

import ibpm
import oracle as O
from cases import boundary_of, case, gaussian_vortex_gamma, oracle_bodies, sc7
from conftest import relerr
from ibpm import _capi as C

pytestmark = pytest.mark.gpu

CD_TOL = 1e-6
NAMES = {1: "FP32-MFMA", 2: "BF16x3-MFMA"}


def set_prec(gp, prec, sweeps=2):
    gp.set_precision({1: "fp32", 2: "bf16x3"}[prec], sweeps=sweeps)
    assert gp.get_option(C.OPT_PRECISION) == prec and gp.get_option(C.OPT_MX_SWEEPS) == sweeps


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("name", ["small3", "spec"])
def test_mixed_lapinv_single_application(name, prec, rng):
    """One application of Δinv / Ainv through the MFMA transforms against the
    oracle: relative error of a single low-precision solve (FP32 about 1e-6,
    split BF16 about 1e-5)."""
    c = case(name)
    g = c["grid"]
    nf = 2 * sum(b.xb.shape[0] for b in c["bodies"])
    op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=np.eye(nf))
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"], Binv=np.eye(nf))
    set_prec(gp, prec)
    tol = 2e-5 if prec == 1 else 2e-4
    for kind in (0, 2):
        for lev in (1, g.mg):
            b = rng.standard_normal(g.nΓ)
            x = np.zeros(g.nΓ)
            gp.ctx.check(C.lib().ibpm_op_lapinv(gp.h, kind, lev, C.ptr(b), C.ptr(x)), "ibpm_op_lapinv")
            xo = op.lapinv(b, kind, lev)
            e = relerr(x, xo)
            print(f"{name} {NAMES[prec]} kind={kind} lev={lev}: {e:.2e}")
            assert e < tol


def run_pair(name, prec, nsteps, sweeps=2, graphs=1, oracle_cache={}):
    """nsteps of the mixed GPU path and of the FP64 oracle (which shares the
    GPU's FP64 Binv) from the same Gaussian-vortex start."""
    c = case(name)
    g = c["grid"]
    key = (name, nsteps)
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    if key not in oracle_cache:
        O.set_threads(16)
        op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), binv=gp.Binv,
                             motion=1 if "motion" in c else 0)
        S = op.new_state(sc7(c, 0.0))
        psi, q, G2 = op.vort2flux(gaussian_vortex_gamma(g))
        S.G[:], S.psi[:], S.q[:] = G2, psi, q
        ts = np.arange(nsteps) * c["dt"]
        cd_o, cl_o = op.run(S, np.stack([sc7(c, t) for t in ts]))
        oracle_cache[key] = (S, cd_o, cl_o)
    S, cd_o, cl_o = oracle_cache[key]
    set_prec(gp, prec, sweeps)
    gp.set_option(C.OPT_GRAPHS, graphs)
    st = ibpm.IBState(gp)
    st.Γ = gaussian_vortex_gamma(g)
    ibpm.vort2flux_(st, gp)
    ts = np.arange(nsteps) * c["dt"]
    cd, cl, done = ibpm.advance_n_(st, gp, ts)
    assert done == nsteps
    scale = float(np.max(np.abs(cd_o)))
    hc2 = (g.h * 2.0 ** np.arange(g.mg)) ** 2
    errs = {
        "CD": float(np.max(np.abs(cd - cd_o)) / scale),
        "CL": float(np.max(np.abs(cl - cl_o)) / scale),
        "CD_final": float(abs(cd[-1, 0] - cd_o[-1, 0]) / abs(cd_o[-1, 0])),
        "ω": relerr(np.asarray(st.Γ) / hc2, S.G / hc2),
        "ψ": relerr(st.ψ, S.psi),
        "q": relerr(st.q, S.q),
    }
    print(f"{name} {NAMES[prec]} sweeps={sweeps} {nsteps} steps: " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    return errs, cd, cl


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("name", ["tiny", "small3", "moving"])
def test_mixed_steps_small(name, prec):
    """Small grids (two or three levels, a static cylinder and the pitching
    plate in the body-fixed frame), 60 steps."""
    errs, _, _ = run_pair(name, prec, 60)
    assert errs["CD"] <= CD_TOL and errs["CL"] <= CD_TOL, errs


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("name", ["cyl_re100_800x400x5", "cyl_re200_1024x1024x5"])
@pytest.mark.timeout(900)
def test_mixed_steps_full_size(name, prec):
    """The benchmark grids (Re 100 cylinder on 800 x 400 with 5 levels, and
    the 5-level 1024 x 1024 Re 200 cylinder), 200 steps from the
    Gaussian-vortex start."""
    errs, _, _ = run_pair(name, prec, 200)
    assert errs["CD"] <= CD_TOL and errs["CL"] <= CD_TOL, errs


def test_mixed_graph_replay_matches_launches():
    """Static-body mixed steps replayed from a captured graph are the same
    launches: bit-identical to stepping without graphs."""
    a, cd_a, cl_a = run_pair("small3", 2, 12, graphs=1)
    b, cd_b, cl_b = run_pair("small3", 2, 12, graphs=0)
    assert np.array_equal(cd_a, cd_b) and np.array_equal(cl_a, cl_b)


def test_one_sweep_is_less_accurate_than_two():
    """The second sweep squares the relative error of each solve."""
    e1, _, _ = run_pair("tiny", 2, 30, sweeps=1)
    e2, _, _ = run_pair("tiny", 2, 30, sweeps=2)
    assert e2["CD"] < 0.1 * e1["CD"]
