"""ibpm_op_lapinv_noalloc: the Laplacian-inverse operator entry that allocates
nothing (fixed per-device scratch), checked against ibpm_op_lapinv, the entry
it is a variant of.  Both run the same kernels, so the results must be
bit-identical.  Two calls in a row, the second on a smaller grid; inputs and
outputs live in device memory the test allocates through HIP itself.
"""
import ctypes

import numpy as np
import pytest

import ibpm
from ibpm import _capi as C

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs for device buffers."""

    def __init__(self):
        # the HIP runtime libibpm_hip.so itself uses (another copy, such as
        # one bundled with a Python package, would be a separate runtime)
        C.lib()
        path = "libamdhip64.so"
        with open("/proc/self/maps") as fh:
            for line in fh:
                f = line.split()
                if f and "libamdhip64" in f[-1] and f[-1].startswith("/opt/rocm"):
                    path = f[-1]
                    break
        self.h = ctypes.CDLL(path)
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, d, a):
        assert self.h.hipMemcpy(d, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device

    def get(self, a, d):
        assert self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), d, a.nbytes, 2) == 0  # device to host


def _problem(name):
    if name == "big":  # 400 x 200, 3 levels: 79,401 points per level
        grid = ibpm.make_grid(0.01, (-1.0, 3.0, -1.0, 1.0), mg=3)
        bodies = [ibpm.make_cylinder(0.25, grid.h, 0.0, 0.0, n=157)]
    else:  # 48 x 32, 2 levels: 1,457 points per level
        grid = ibpm.make_grid(0.125, (-2.0, 4.0, -2.0, 2.0), mg=2)
        bodies = [ibpm.make_cylinder(0.5, grid.h, 0.0, 0.1, n=24)]
    nf = 2 * sum(b.xb.shape[0] for b in bodies)
    return ibpm.IBProblem(grid, bodies, 0.002, 100.0, freestream={"Ux": lambda t: 1.0}, Binv=np.eye(nf))


def test_noalloc_matches_lapinv_twice_second_smaller():
    hip = _Hip()
    L = C.lib()
    rng = np.random.default_rng(7)
    for name in ("big", "small"):  # the second call has the smaller input
        prob = _problem(name)
        n = int(prob.info.nG)
        assert n <= C.IBPM_NOALLOC_MAX_POINTS
        d_b, d_x = hip.alloc(8 * n), hip.alloc(8 * n)
        try:
            for kind in (0, 1, 2):
                for lev in sorted({1, prob.info.mg}):
                    b = rng.standard_normal(n)
                    ref = np.zeros(n)
                    prob.ctx.check(L.ibpm_op_lapinv(prob.h, kind, lev, C.ptr(b), C.ptr(ref)), "ibpm_op_lapinv")
                    hip.put(d_b, b)
                    hip.put(d_x, np.full(n, np.nan))
                    prob.ctx.check(L.ibpm_op_lapinv_noalloc(prob.h, kind, lev, d_b, d_x, n),
                                   "ibpm_op_lapinv_noalloc")
                    got = np.zeros(n)
                    hip.get(got, d_x)
                    print(f"{name} n={n} kind={kind} lev={lev}: max |diff| = {np.max(np.abs(got - ref)):.3e}",
                          flush=True)
                    assert np.array_equal(got, ref)
            # a size that is not the problem's level size is refused
            rc = L.ibpm_op_lapinv_noalloc(prob.h, 0, 1, d_b, d_x, n + 1)
            assert rc == C.IBPM_E_ARG
        finally:
            hip.h.hipFree(d_b)
            hip.h.hipFree(d_x)
