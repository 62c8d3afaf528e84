"""Write the reference-data fixtures of tests/golden from the reference tree.

Run in a container that has the reference mounted:
    python tests/golden/make_golden.py /root/reference

Outputs (data only -- no reference source is copied):
  eldredge_force_jl2.npz  t, CL, CD of benchmarks/eldredge/results/force_jl2.mat,
                          the Julia body-fixed (MovingGrid) pitch-hold run
                          (benchmarks/eldredge/pitch_maneuver.jl:7-78), lab-rotated.
                          MAT v7.3 = HDF5; the three 10001-double datasets are
                          stored contiguously and uncompressed at the byte
                          offsets below (checked: t runs 0:0.001:10).
  eldredge_force_jl1.npz  t, CL, CD of force_jl1.mat, the Julia lab-frame
                          (MotionFunction) run of the same maneuver
                          (pitch_maneuver_motion.jl): raw state.CD/CL, which
                          are lab-frame forces there; same layout and offsets.
                          Its last sample is 0 (the animation loop runs
                          100 x 100 of the 10001 steps).
  eldredge_force_fort.npz step, CD, CL columns (4, 5) of force_fort.dat, the
                          Fortran run of the same maneuver (plot_results.jl:46-49).
  reference_numbers.json  published scalar results with their source lines.
"""
import json
import os
import sys

import numpy as np

OFFSETS = {"t": 2560, "CL": 82568, "CD": 162576}
N = 10001


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    res = os.path.join(ref, "benchmarks", "eldredge", "results")
    raw = open(os.path.join(res, "force_jl2.mat"), "rb").read()
    arrs = {k: np.frombuffer(raw, dtype="<f8", count=N, offset=o).copy() for k, o in OFFSETS.items()}
    assert np.allclose(arrs["t"], np.arange(N) * 1e-3, atol=1e-12)
    np.savez_compressed(os.path.join(here, "eldredge_force_jl2.npz"), **arrs)
    raw1 = open(os.path.join(res, "force_jl1.mat"), "rb").read()
    arrs1 = {k: np.frombuffer(raw1, dtype="<f8", count=N, offset=o).copy() for k, o in OFFSETS.items()}
    assert np.allclose(arrs1["t"], np.arange(N) * 1e-3, atol=1e-12)
    np.savez_compressed(os.path.join(here, "eldredge_force_jl1.npz"), **arrs1)
    fort = np.loadtxt(os.path.join(res, "force_fort.dat"))
    np.savez_compressed(os.path.join(here, "eldredge_force_fort.npz"), step=fort[:, 0], CD=fort[:, 3], CL=fort[:, 4])
    numbers = {
        "steady_flat_plate_CD_julia": {"value": 0.39415052, "source": "benchmarks/eldredge/README.md:25",
                                       "case": "benchmarks/eldredge/steady_bl.jl + config.jl: 400x200, mg=5, Re=100, "
                                               "plate L=1 at 0 deg with nb=51, dt=0.01, t=0:0.01:100"},
        "steady_flat_plate_CD_fortran": {"value": 0.39415060, "source": "benchmarks/eldredge/README.md:26"},
        "cyl40_multilevel_CD_julia": {"value": 1.5377938, "source": "benchmarks/README.md:14"},
        "cyl40_single_level_CD_julia": {"value": 2.5060811090446, "source": "benchmarks/README.md:20",
                                        "case": "benchmarks/cyl_40.jl (mg=1, 200x200, Re=40, dt=0.01); the final "
                                                "time of the tabulated run is not stated"},
        "cyl40_single_level_CD_fortran": {"value": 2.5060811857143, "source": "benchmarks/README.md:19"},
    }
    with open(os.path.join(here, "reference_numbers.json"), "w") as f:
        json.dump(numbers, f, indent=1)
    print("wrote fixtures to", here)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
