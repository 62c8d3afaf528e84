"""Full runs of the benchmark configurations against the oracle's histories.

tests/golden/make_long_histories.py ran the CPU oracle over each
configuration's full step count (its own B, assembled and inverted by
get_Binv) and kept CD and CL of every step and the norms of Γ and ψ per
level every 500 steps.  The GPU runs the same configuration from the same
start through ibpm_advance_n with the B it assembles and inverts itself, and
is held to the oracle over the whole history:

  * cyl_re40_200x200x3: Re 40 cylinder, impulsive start, 1000 steps;
  * cyl_re100_800x400x5: Re 100 cylinder, Gaussian-vortex start, 5000 steps;
  * cyl_re200_1024x1024x5: Re 200 cylinder, Gaussian-vortex start, the first
    2000 steps of the 10000.

Tolerances: CD to 1e-10 relative (every step), CL to 1e-10 of max |CD|, the
Γ and ψ norms of every level to 1e-10 relative.  Measured on MI355X: CD and
CL within 4e-14 on all three runs, the norms within 3e-13 (200 x 200),
5.5e-12 (800 x 400 after 5000 steps) and 2.8e-11 (ψ at 1024 x 1024 after
2000 steps): two FP64 solvers whose transforms round differently drift apart
slowly, and a wrong operator moves CD by far more than 1e-10 within a few
steps.
"""
import os

import numpy as np
import pytest

import ibpm
from cases import case, gaussian_vortex_gamma

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10

CASES = [("cyl_re40_200x200x3", False), ("cyl_re100_800x400x5", True), ("cyl_re200_1024x1024x5", True)]


@pytest.mark.parametrize("name,vortex", CASES)
def test_full_run_matches_oracle_history(name, vortex):
    path = os.path.join(GOLDEN, f"oracle_long_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"tests/golden/oracle_long_{name}.npz not generated")
    d = np.load(path)
    c = case(name)
    g = c["grid"]
    gp = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
    st = ibpm.IBState(gp)
    if vortex:
        st.Γ = gaussian_vortex_gamma(g)
        ibpm.vort2flux_(st, gp)
    n = len(d["CD"])
    cd = np.zeros(n)
    cl = np.zeros(n)
    nG, npsi = [], []
    done = 0
    for stop in d["norm_steps"]:
        ts = np.arange(done, stop) * c["dt"]
        a, b, k = ibpm.advance_n_(st, gp, ts)
        assert k == len(ts)
        cd[done:stop], cl[done:stop] = a[:, 0], b[:, 0]
        nG.append(np.linalg.norm(st.Γ, axis=0))
        npsi.append(np.linalg.norm(st.ψ, axis=0))
        done = stop
    ecd = np.max(np.abs(cd - d["CD"])) / np.max(np.abs(d["CD"]))
    ecl = np.max(np.abs(cl - d["CL"])) / np.max(np.abs(d["CD"]))
    eG = np.max(np.abs(np.array(nG) - d["norm_G"]) / d["norm_G"])
    epsi = np.max(np.abs(np.array(npsi) - d["norm_psi"]) / d["norm_psi"])
    print(f"{name}, {n} steps: CD {ecd:.2e}, CL {ecl:.2e} (of max |CD|), |Γ| {eG:.2e}, |ψ| {epsi:.2e}; "
          f"final CD {cd[-1]:.12f} (oracle {d['CD'][-1]:.12f})")
    assert ecd < TOL and ecl < TOL
    assert eG < TOL and epsi < TOL
