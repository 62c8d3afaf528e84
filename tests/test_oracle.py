"""Pins of the CPU oracle (CPU only).

The reference has no automated tests and no vectors for individual operators,
so the oracle is pinned here by what the mathematics fixes, each check chosen
so that a plausible slip (a dropped term, a wrong sign or index, a transposed
operand) fails it:
  * DST-I against the brute-force O(n^2) sine sum and against scipy's
    independent RODFT00 (scipy.fft.dstn type 1), and its involution;
  * Δinv / A / Ainv against the assembled 5-point operators C'C they invert
    (laplacian.jl:14-76 states Λ are the eigenvalues of RC);
  * curl/rot and E/E' adjointness (curl.jl:22-71, interface-coupling.jl:99-133);
  * Yang-3 δh moments, support and symmetry (interface-coupling.jl:6-62);
  * coarsify!/get_bc!/apply_bc! on fields where they reduce to closed forms;
  * zero-vorticity nonlinear term and the uniform base flux.
End-to-end pins against the Julia-produced numbers in the reference tree are
in test_reference_data.py.
"""
import math

import numpy as np
import pytest
import scipy.fft as sfft

import ibpm
import oracle as O
from cases import boundary_of, case, oracle_bodies, sc7
from conftest import relerr


def brute_dst1(x):
    n = len(x)
    k = np.arange(1, n + 1)
    S = 2.0 * np.sin(np.pi * np.outer(k, k) / (n + 1))
    return S @ x


@pytest.mark.parametrize("n1,n2", [(1, 1), (1, 7), (5, 3), (8, 9), (15, 4), (24, 11)])
def test_dst2_brute_force(rng, n1, n2):
    x = rng.standard_normal((n1, n2))
    # separable 2D transform from the brute-force 1D sum
    y = np.apply_along_axis(brute_dst1, 0, x)
    y = np.apply_along_axis(brute_dst1, 1, y)
    assert relerr(O.dst2(x), y) < 1e-13


@pytest.mark.parametrize("n1,n2", [(199, 199), (399, 199), (127, 255), (119, 39)])
def test_dst2_against_scipy(rng, n1, n2):
    x = rng.standard_normal((n1, n2))
    assert relerr(O.dst2(x), sfft.dstn(x, type=1, axes=[0, 1])) < 1e-14


def test_dst2_involution(rng):
    x = rng.standard_normal((63, 31))
    y = O.dst2(O.dst2(x)) / (4.0 * 64 * 32)
    assert relerr(y, x) < 1e-14


@pytest.fixture(scope="module")
def small():
    c = case("small3")
    g = c["grid"]
    P = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c))
    return c, P


def laplacian_5pt(P, psi):
    """rot(curl(ψ)) on one grid with zero boundary values, i.e. -Δ_h ψ (h = 1)."""
    return P.rot(P.curl_sg(psi))


def test_rot_curl_is_five_point(small, rng):
    _, P = small
    n1, n2 = P.nx - 1, P.ny - 1
    psi = rng.standard_normal((n1, n2))
    pad = np.zeros((n1 + 2, n2 + 2))
    pad[1:-1, 1:-1] = psi
    lap = 4 * psi - pad[2:, 1:-1] - pad[:-2, 1:-1] - pad[1:-1, 2:] - pad[1:-1, :-2]
    assert relerr(laplacian_5pt(P, psi.ravel(order="F")), lap.ravel(order="F")) < 1e-14


def test_lap_eigs_closed_form(small):
    _, P = small
    i = np.arange(1, P.nx)[:, None]
    j = np.arange(1, P.ny)[None, :]
    lam = 4 - 2 * np.cos(np.pi * i / P.nx) - 2 * np.cos(np.pi * j / P.ny)
    assert relerr(P.lam(), lam) < 1e-15


def test_dinv_inverts_rot_curl(small, rng):
    _, P = small
    psi = rng.standard_normal(P.nG)
    assert relerr(P.lapinv(laplacian_5pt(P, psi), 0), psi) < 1e-12


@pytest.mark.parametrize("lev", [1, 2, 3])
def test_ainv_and_a_match_stencils(small, rng, lev):
    c, P = small
    hc = P.h * 2 ** (lev - 1)
    k = c["dt"] / (2 * c["Re"] * hc * hc)
    x = rng.standard_normal(P.nG)
    # Ainv solves (I + k C'C) y = x  (laplacian.jl:61-71)
    y = P.lapinv(x, 2, lev)
    assert relerr(y + k * laplacian_5pt(P, y), x) < 1e-13
    # A applies (I - k C'C)  (laplacian.jl:73-76, 78-81)
    assert relerr(P.lapinv(x, 1, lev), x - k * laplacian_5pt(P, x)) < 1e-12


def test_curl_rot_adjoint(small, rng):
    _, P = small
    psi = rng.standard_normal(P.nG)
    q = rng.standard_normal(P.nq)
    assert abs(P.curl_sg(psi) @ q - psi @ P.rot(q)) < 1e-10 * np.linalg.norm(psi) * np.linalg.norm(q)


def test_E_Et_adjoint(small, rng):
    _, P = small
    q = rng.standard_normal(P.nq)
    f = rng.standard_normal(2 * P.nbt)
    assert abs(P.regT(q) @ f - q @ P.reg(f)) < 1e-12 * np.linalg.norm(q) * np.linalg.norm(f)


def test_E_of_uniform_flux(small):
    _, P = small
    q = np.zeros(P.nq)
    q[: P.nu] = 0.7
    q[P.nu:] = -0.3
    fb = P.regT(q)
    # Σ weights = (Σ δh)^2 = 1 to the 10-digit rounding of the Yang-3 constants
    assert np.max(np.abs(fb[: P.nbt] - 0.7)) < 1e-8
    assert np.max(np.abs(fb[P.nbt:] + 0.3)) < 1e-8


def test_delta_h_moments():
    h = 1.0
    for off in np.linspace(0.0, 1.0, 37):
        pts = np.arange(-6, 7) + off
        w = np.array([O.delta_h(p, 0.0, h) for p in pts])
        # the reference's constants carry 10 significant digits, so the
        # moments hold to a few 1e-9 (δh(3h) itself is -3.4e-9); a wrong
        # coefficient or branch breaks them at 1e-3 or worse
        assert abs(w.sum() - 1.0) < 1e-8
        assert abs((w * pts).sum()) < 1e-8
        assert np.all(w[np.abs(pts) > 3.0] == 0.0)
    # symmetric in its argument, continuous across the branch points r = h, 2h
    assert O.delta_h(0.7, 0.0, 1.0) == O.delta_h(-0.7, 0.0, 1.0)
    for rb in (1.0, 2.0):
        assert abs(O.delta_h(rb - 1e-12, 0.0, 1.0) - O.delta_h(rb + 1e-12, 0.0, 1.0)) < 1e-8


def test_body_idx_and_window(small):
    c, P = small
    g = c["grid"]
    xb = c["bodies"][0].xb
    bidx = P.body_idx()
    assert np.array_equal(bidx[:, 0], np.floor((xb[:, 0] + g.offx) / g.h).astype(int))
    assert np.array_equal(bidx[:, 1], np.floor((xb[:, 1] + g.offy) / g.h).astype(int))
    W = P.weights()
    # only a 6 x 6 (or 7 x 7 at exact alignment) block of the 13 x 13 window is non-zero
    assert np.all(np.count_nonzero(W[:, 0], axis=(1, 2)) <= 49)


def test_coarsify_closed_form(small, rng):
    _, P = small
    nx, ny = P.nx, P.ny
    i = np.arange(1, nx)[:, None]
    j = np.arange(1, ny)[None, :]
    a, b, c0 = 0.3, -1.1, 2.0
    G = (a * i + b * j + c0).ravel(order="F").astype(np.float64)
    Gc = rng.standard_normal(P.nG)
    out = P.coarsify(Gc, G).reshape(nx - 1, ny - 1, order="F")
    Gc2 = Gc.reshape(nx - 1, ny - 1, order="F")
    # centre block: 4 x (linear field at the fine point 2ic - nx/2); ring untouched
    ic = np.arange(nx // 2 - nx // 4 + 1, nx // 2 + nx // 4)
    jc = np.arange(ny // 2 - ny // 4 + 1, ny // 2 + ny // 4)
    fi = 2 * ic - nx // 2
    fj = 2 * jc - ny // 2
    expect = 4 * (a * fi[:, None] + b * fj[None, :] + c0)
    assert relerr(out[np.ix_(ic - 1, jc - 1)], expect) < 1e-14
    mask = np.ones_like(out, dtype=bool)
    mask[np.ix_(ic - 1, jc - 1)] = False
    assert np.array_equal(out[mask], Gc2[mask])


def test_get_bc_apply_bc_linear(small):
    _, P = small
    nx, ny = P.nx, P.ny
    i = np.arange(1, nx)[:, None]
    j = np.arange(1, ny)[None, :]
    # coarse field linear in the coarse index; the fine boundary is at coarse
    # rows/cols nx/4, 3nx/4 and interpolates linearly in between
    r = (0.5 * i - 0.25 * j).ravel(order="F").astype(np.float64)
    bc = P.get_bc(r)
    B, T, L, R = 2 * (ny + 1), 2 * (ny + 1) + nx + 1, 0, ny + 1
    ib = np.arange(1, nx + 2)  # fine boundary positions
    ci = nx // 4 + (ib - 1) / 2.0
    assert relerr(bc[B:B + nx + 1], 0.5 * ci - 0.25 * (ny // 4)) < 1e-15
    assert relerr(bc[T:T + nx + 1], 0.5 * ci - 0.25 * (3 * ny // 4)) < 1e-15
    jb = np.arange(1, ny + 2)
    cj = ny // 4 + (jb - 1) / 2.0
    assert relerr(bc[L:L + ny + 1], 0.5 * (nx // 4) - 0.25 * cj) < 1e-15
    assert relerr(bc[R:R + ny + 1], 0.5 * (3 * nx // 4) - 0.25 * cj) < 1e-15
    # apply_bc!: only the first/last interior rows and columns change
    z = P.apply_bc(np.zeros(P.nG), bc, 2.0).reshape(nx - 1, ny - 1, order="F")
    assert np.all(z[1:-1, 1:-1] == 0.0)
    # r[i, 1] += fac * rbc[(B+1)+i] (1-based) is bc[B + i] here (0-based)
    assert z[3, 0] == 2.0 * bc[B + 4]
    assert z[0, 0] == 2.0 * bc[B + 1] + 2.0 * bc[L + 1]
    assert z[nx - 2, ny - 2] == 2.0 * bc[T + nx - 1] + 2.0 * bc[R + ny - 1]


def test_nonlinear_zero_vorticity(small, rng):
    c, P = small
    q = np.asfortranarray(rng.standard_normal((P.nq, P.mg)))
    q0 = np.asfortranarray(rng.standard_normal((P.nq, P.mg)))
    G = np.zeros((P.nG, P.mg), order="F")
    for lev in range(1, P.mg + 1):
        assert np.all(P.nonlinear(q, q0, G, lev) == 0.0)


def test_nonlinear_uniform_flow_uniform_vorticity(small):
    # uniform flux and uniform Γ: the divergence of Q Γ vanishes in the interior
    c, P = small
    q = np.zeros((P.nq, P.mg), order="F")
    q0 = np.zeros((P.nq, P.mg), order="F")
    q0[: P.nu, :] = 1.0
    G = np.ones((P.nG, P.mg), order="F")
    out = P.nonlinear(q, q0, G, 1).reshape(P.nx - 1, P.ny - 1, order="F")
    assert np.max(np.abs(out[1:-1, 1:-1])) < 1e-12
    assert np.max(np.abs(out)) > 0.0  # the boundary rows see Γbc from level 2


def test_base_flux_uniform(small):
    c, P = small
    S = P.new_state(np.array([1.0, 0.5, 0.2, 0, 0, 0, 0]))
    for lev in range(1, P.mg + 1):
        hc = P.h * 2 ** (lev - 1)
        assert np.allclose(S.q0[: P.nu, lev - 1], (1.0 * math.cos(0.2) - 0.5 * math.sin(0.2)) * hc, rtol=0, atol=0)
        assert np.allclose(S.q0[P.nu:, lev - 1], (1.0 * math.sin(0.2) + 0.5 * math.cos(0.2)) * hc, rtol=0, atol=0)


def test_vort2flux_single_level_inverse(rng):
    # on one level with zero boundary data, q = C ψ and rot(q) = Γ exactly
    c = case("cyl40_single")
    g = c["grid"]
    P = O.OracleProblem(g.h, boundary_of(g), 1, c["Re"], c["dt"], oracle_bodies(c), binv=np.eye(2 * 78))
    G = rng.standard_normal((P.nG, 1))
    psi, q, _ = P.vort2flux(G)
    assert relerr(P.rot(q[:, 0]), G[:, 0]) < 1e-12
    assert relerr(laplacian_5pt(P, psi[:, 0]), G[:, 0]) < 1e-12


def test_invert_matches_numpy(rng):
    A = rng.standard_normal((50, 50)) + 10 * np.eye(50)
    assert relerr(O.invert(A), np.linalg.inv(A)) < 1e-13


def test_binv_inverts_B(small):
    c, P = small
    nf = 2 * P.nbt
    B = np.column_stack([P.B_times(np.eye(nf)[:, j]) for j in range(nf)])
    assert relerr(P.binv() @ B, np.eye(nf)) < 1e-9


def test_step_conserves_zero_state_without_flow(small):
    # no freestream, zero state: advance! keeps everything zero
    c, P = small
    S = P.new_state(np.zeros(7))
    P.advance(S, np.zeros(7))
    assert np.all(S.G == 0) and np.all(S.q == 0) and np.all(S.CD == 0)


# ---------------------------------------------------------------------------
# LinearizedIBProblem (problem-types.jl:47-87, nonlinear.jl:45-65,144-163)
# ---------------------------------------------------------------------------
def _lin_pair(mg=1, seed=3):
    """A problem and its linearization about a random base state."""
    grid = ibpm.make_grid(0.125, (-2.0, 4.0, -2.0, 2.0), mg=mg)
    bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.1, n=24)]
    args = (grid.h, boundary_of(grid), mg, 40.0, 0.02, [(b.xb, b.ds) for b in bodies])
    eye = np.eye(2 * 24)
    full = O.OracleProblem(*args, binv=eye)
    lin = O.OracleProblem(*args, binv=eye)
    rng = np.random.default_rng(seed)
    qB = np.asfortranarray(rng.standard_normal((grid.nq, mg)))
    q0B = np.asfortranarray(np.full((grid.nq, mg), 0.3))
    GB = np.asfortranarray(rng.standard_normal((grid.nΓ, mg)))
    lin.linearize(qB, q0B, GB)
    return grid, full, lin, qB, q0B, GB, rng


def test_linearized_nonlinear_is_the_cross_term_of_the_bilinear_product():
    # one level: no boundary ring, so N(Q, Γ) = rot(Q x Γ)/(4h^2) is bilinear
    grid, full, lin, qB, q0B, GB, rng = _lin_pair()
    q = np.asfortranarray(rng.standard_normal((grid.nq, 1)))
    G = np.asfortranarray(rng.standard_normal((grid.nΓ, 1)))
    z = np.zeros_like(q)
    tot = full.nonlinear(qB + q, q0B, GB + G, 1)
    base = full.nonlinear(qB, q0B, GB, 1)
    pert = full.nonlinear(q, z, G, 1)
    cross = lin.nonlinear(q, z, G, 1)
    scale = max(np.abs(tot).max(), np.abs(base).max(), np.abs(pert).max())
    assert np.abs(cross - (tot - base - pert)).max() < 1e-12 * scale
    assert np.abs(cross).max() > 1e-3 * scale  # not trivially zero


def test_linearized_nonlinear_homogeneous_and_zero_base():
    grid, full, lin, qB, q0B, GB, rng = _lin_pair(mg=2)
    q = np.asfortranarray(rng.standard_normal((grid.nq, 2)))
    G = np.asfortranarray(rng.standard_normal((grid.nΓ, 2)))
    z = np.zeros_like(q)
    # Homogeneous wherever the boundary ring is zero: on the coarsest level, and
    # on level 1 when the coarse level of the perturbation is zero.  With a ring,
    # the reference multiplies the perturbation's Q by the perturbation's own
    # Γbc in the Q x ΓB product (direct_product!, nonlinear.jl:155-157), which
    # is quadratic -- a quirk reproduced here.
    Gz = G.copy()
    Gz[:, 1] = 0.0
    for lev, GG in ((2, G), (1, Gz)):
        a = lin.nonlinear(q, z, GG, lev)
        b = lin.nonlinear(2 * q, z, 2 * GG, lev)
        assert np.array_equal(b, 2 * a)  # every operation is linear; scaling by 2 is exact
    a = lin.nonlinear(q, z, G, 1)
    b = lin.nonlinear(2 * q, z, 2 * G, 1)
    assert not np.allclose(b, 2 * a)  # the quadratic boundary term
    lin0 = O.OracleProblem(grid.h, boundary_of(grid), 2, 40.0, 0.02, [(np.zeros((1, 2)), np.ones(1))],
                           binv=np.eye(2))
    lin0.linearize(np.zeros_like(q), np.zeros_like(q), np.zeros_like(G))
    G0 = G.copy()
    G0[:, 1] = 0.0  # no boundary ring from the coarse level either
    assert np.array_equal(lin0.nonlinear(q, z, G0, 1), np.zeros(grid.nΓ))
    S = lin.new_state(np.array([1.0, 0.0, 0.0, 0, 0, 0, 0]))
    assert np.all(S.q0 == 0.0)  # base_flux! of the linearized problem (timestepping.jl:238-250)
