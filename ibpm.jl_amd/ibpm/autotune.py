"""Launch-shape autotuner for a problem.

Each option is swept over its legal values with the kernel it controls timed by
HIP events on the problem's stream (ibpm_time_kernel); the fastest value is
kept.  Results depend only on the grid shape, so one sweep per problem
suffices.
"""
from __future__ import annotations

import ctypes

from . import _capi as C

# option -> (kernel id timed by ibpm_time_kernel, candidate values).  The
# one-level y-passes are mostly those of the vort2flux! chain, which carry the
# boundary terms (id 8); a one-level problem has none and times id 5.
SWEEPS = {
    C.OPT_YTILE_BATCH: (1, (4, 8, 12)),
    C.OPT_YTILE_SINGLE: (8, (4, 8, 12)),
    C.OPT_XWAVES_BATCH: (0, (4, 8, 12, 14)),
    # persistent prefetching work-groups for the passes over many tiles
    # (timed on the all-level y-pass, after its tile is chosen)
    C.OPT_PIPE: (1, (1, 0)),
    # two wavefronts per column in the one-level y-passes of 1023-point
    # columns (no effect on other grids: both values time the same kernel)
    C.OPT_YSPLIT: (8, (1, 0)),
}


def time_kernel(prob, which: int, reps: int = 20) -> float:
    v = ctypes.c_double(0.0)
    prob.ctx.check(C.lib().ibpm_time_kernel(prob.h, which, reps, ctypes.byref(v)), "ibpm_time_kernel")
    return v.value


def autotune(prob, reps: int = 20) -> dict:
    """Sweep every option of SWEEPS on `prob`, apply the fastest values and
    return {option: (best_value, {value: ms})}."""
    L = C.lib()
    out = {}
    for opt, (which, values) in SWEEPS.items():
        if which == 8 and prob.model.grid.mg < 2:
            which = 5
        times = {}
        for v in values:
            if L.ibpm_set_option(prob.h, opt, v) != C.IBPM_OK:
                continue  # does not fit this grid
            times[v] = time_kernel(prob, which, reps)
        best = min(times, key=times.get)
        prob.ctx.check(L.ibpm_set_option(prob.h, opt, best), "ibpm_set_option")
        out[opt] = (best, times)
    return out


# options of a group's all-level passes, swept by timing whole group steps:
# the per-problem sweep times one instance's launch, while a group launches
# every instance at once and may prefer other shapes
GROUP_SWEEPS = (
    ((C.OPT_YTILE_BATCH, (4, 8, 12)), (C.OPT_PIPE, (1, 0))),
    ((C.OPT_XWAVES_BATCH, (4, 8, 12, 14)),),
)


def autotune_group(group, steps: int = 6) -> dict:
    """Sweep the all-level launch shapes of `group` (the y tile with the
    persistent work-groups, then the x-pass work-group size) by the wall time
    of `steps` group steps per candidate, best of two; apply the fastest and
    return {option: best value}.  The timed steps advance the group's state,
    as warm-up steps would; every shape gives the same results."""
    import itertools
    import time

    def step_us():
        best = float("inf")
        for _ in range(2):
            t = time.perf_counter()
            group.advance_async(steps)
            group.wait()
            best = min(best, (time.perf_counter() - t) / steps * 1e6)
        return best

    chosen = {}
    for sweep in GROUP_SWEEPS:
        opts = [o for o, _ in sweep]
        times = {}
        for combo in itertools.product(*[vals for _, vals in sweep]):
            try:
                for o, v in zip(opts, combo):
                    group.set_option(o, v)
            except C.IBPMError:
                continue  # does not fit this grid
            group.advance_async(1)
            group.wait()
            times[combo] = step_us()
        best = min(times, key=times.get)
        for o, v in zip(opts, best):
            group.set_option(o, v)
            chosen[o] = v
    return chosen
