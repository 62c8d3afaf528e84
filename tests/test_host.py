"""Host-side logic and the C ABI surface (CPU only: no compute calls)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import ibpm
import oracle as O
from cases import boundary_of, case, oracle_bodies
from ibpm import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ibpm_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ibpm_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 20
    assert sorted(_capi.EXPORTED) == names
    L = ctypes.CDLL(_capi.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_library_is_gfx950_code_object():
    blob = open(_capi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


@pytest.mark.parametrize("name", ["tiny", "small3", "mixed", "moving", "spec", "cyl_re40_200x200x3", "steady_bl", "cyl_re100_800x400x5", "cyl_re200_1024x1024x5"])
def test_make_grid_matches_oracle(name):
    c = case(name)
    g = c["grid"]
    P = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c),
                        binv=np.eye(sum(2 * len(b.ds) for b in c["bodies"])), motion=1 if "motion" in c else 0)
    assert (P.nx, P.ny, P.nG, P.nq, P.mg) == (g.nx, g.ny, g.nΓ, g.nq, g.mg)
    assert (P.h, P.offx, P.offy) == (g.h, g.offx, g.offy)


def test_make_grid_fields():
    g = ibpm.make_grid(0.02, (-1.5, 6.5, -2.0, 2.0), mg=5)
    assert (g.nx, g.ny, g.mg) == (400, 200, 5)
    assert (g.LEFT, g.RIGHT, g.BOT, g.TOP) == (0, 201, 402, 803)
    assert g.nq == 200 * 401 + 400 * 201
    q = np.arange(g.nq, dtype=float)
    qx, qy = g.split_flux(q)
    assert qx.shape == (401, 200) and qy.shape == (400, 201)
    assert qx[1, 0] == 1.0 and qy[0, 0] == g.nu


def test_make_cylinder_closed_form():
    b = ibpm.make_cylinder(0.5, 0.02, 1.0, -0.5)
    n = int(math.floor(2 * math.pi * 0.5 / 0.02 / 2))
    assert b.xb.shape == (n, 2) == (78, 2)
    assert np.allclose(np.hypot(b.xb[:, 0] - 1.0, b.xb[:, 1] + 0.5), 0.5, atol=1e-15)
    assert np.allclose(b.ds, 2 * 0.5 * math.sin(math.pi / n), rtol=1e-14)
    assert isinstance(b.motion, ibpm.Static)


def test_make_plate_closed_form():
    b = ibpm.make_plate(1.0, 0.2, 0.02, 0.0, 0.0, n=51)
    assert b.xb.shape == (51, 2)
    assert np.allclose(b.xb[-1], [math.cos(0.2), -math.sin(0.2)], atol=1e-15)
    assert np.allclose(b.ds, 1.0 / 50, rtol=1e-14)
    b2 = ibpm.make_plate(1.0, 0.0, 0.02, 0.0, 0.0)
    assert b2.xb.shape[0] == 26  # floor(L / h / 2) + 1


def test_naca_symmetric_thickness():
    _, xU, xL, yU, yL = ibpm.naca(np.array([0.0, 0.3, 1.0]), "0012")
    assert np.allclose(yU, -yL)
    assert abs(yU[1] - 0.06) < 2e-4  # max thickness 12% of chord near x = 0.3
    b = ibpm.make_naca(0.25, 40, "0012")
    assert b.xb.shape == (78, 2) and np.all(b.ds > 0)


def test_motion_type_dispatch():
    st = ibpm.make_cylinder(0.5, 0.02, 0, 0)
    mg = ibpm.make_plate(1.0, 0.0, 0.02, 0, 0, motion=ibpm.MovingGrid(lambda t: 1, lambda t: 0, lambda t: 0, lambda t: 0))
    assert ibpm.MotionType([st, st]) is ibpm.Static
    assert ibpm.MotionType([mg]) is ibpm.MovingGrid
    assert ibpm.MotionType([st, mg]) is ibpm.Motion
    assert ibpm.get_body_info([st, mg]) == ([78, 26], [156, 52])


def test_make_body_specs():
    bodies = ibpm.make_body([{"type": "cylinder", "lengthscale": 0.5, "center": (0.0, 0.1)},
                             {"type": "plate", "lengthscale": 1.0, "spec": {"α": 0.1}, "motion": "static"}], 0.02)
    assert len(bodies) == 2 and bodies[0].xb.shape[0] == 78


def test_step_scalars_layout():
    s = _capi.StepScalars(1, 2, 3, 4, 5, 6, 7)
    assert ctypes.sizeof(s) == 7 * 8 + 8  # seven doubles, then the body_kin pointer
    assert _capi.StepScalars.body_kin.offset == 7 * 8 and not s.body_kin
    assert ctypes.sizeof(_capi.GridDesc) == 3 * 4 + 4 + 7 * 8  # padding after mg


def test_init_save_data_default_keeps_final_state():
    t = np.arange(10) * 0.01
    data = ibpm.init_save_data(t, None, None)
    assert len(data) == 1 and np.allclose(data[0].t, [0.09])


def test_init_save_data_times():
    t = np.arange(101) * 0.01
    f = lambda tt, st: tt
    data = ibpm.init_save_data(t, {"save_fcns": [f, f, f, f], "save_times": [None, 0.0, 0.05, [0.031, 0.5, 0.5004]]}, None)
    assert len(data[0].t) == 101 and len(data[1].t) == 101
    assert np.allclose(data[2].t, np.arange(21) * 0.05)          # every round(0.05/dt) dt from t[0]
    assert np.allclose(data[3].t, [0.03, 0.5])                   # nearest step times, unique
    from ibpm.problem import _save_steps
    assert _save_steps(t, [data[2]]) == set(range(0, 101, 5))
    assert _save_steps(t, [data[3]]) == {3, 50}
    ibpm.save_data_(0.5, t, None, None, data)
    assert [len(d.val) for d in data] == [1, 1, 1, 1]
    ibpm.save_data_(0.51, t, None, None, data)
    assert [len(d.val) for d in data] == [2, 2, 1, 1]
