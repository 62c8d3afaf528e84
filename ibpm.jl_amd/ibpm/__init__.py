"""ibpm -- MI355X (gfx950) drop-in for IBPM.jl's `advance!` hot path.

Python mirror of the reference's Julia module `ibpm` (src/ibpm.jl) for the
path it accelerates.  Julia's mutating `name!` functions are spelled `name_`.

    import ibpm
    grid = ibpm.make_grid(0.02, (-1.0, 3.0, -2.0, 2.0), mg=3)
    body = ibpm.make_cylinder(0.5, grid.h, 0.0, 0.0, n=157)
    prob = ibpm.IBProblem(grid, [body], 0.004, 40.0, freestream=dict(Ux=lambda t: 1.0))
    state = ibpm.IBState(prob)
    ibpm.advance_(state, prob, 0.0)      # advance!(state, prob, t)
    state.CD, state.CL

All computation runs in libibpm_hip.so on the GPU; importing works without a
GPU, constructing an IBProblem needs one.
"""
from .domain import (InertialMotion, Motion, MotionFunction, MotionType, MovingGrid, MultiGrid, RigidBody, Static,
                     get_body_info, make_body, make_cylinder, make_grid, make_naca, make_plate, naca)
from .problem import (AB2, AdamsBashforth, Group, IBModel, IBProblem, IBState, LinearizedIBProblem, StateSnapshot,
                      UserVar,
                      advance, advance_, advance_batch_, advance_groups_,
                      advance_n_, base_flux_, compute_cfl, copy_, init_save_data, load_state, run_sim_, save_data_,
                      save_state, snapshot, vort2flux_)
from . import _capi, adjoint

__all__ = [
    "Group",
    "MultiGrid", "make_grid", "Motion", "InertialMotion", "Static", "MovingGrid", "MotionFunction", "RigidBody",
    "make_plate", "make_cylinder", "make_naca", "naca", "make_body", "MotionType", "get_body_info", "AB2",
    "AdamsBashforth", "IBModel", "IBProblem", "LinearizedIBProblem", "IBState", "advance_", "advance", "advance_n_", "advance_batch_", "advance_groups_", "base_flux_",
    "vort2flux_", "compute_cfl", "run_sim_", "UserVar", "init_save_data", "save_data_", "StateSnapshot", "snapshot",
    "copy_", "save_state", "load_state", "adjoint",
]
