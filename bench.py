"""Benchmark: timesteps/s of the FP64 advance! step on one MI355X.

Workload: impulsively started cylinder (diameter 1) at Re = 200 in uniform
flow, finest grid 1024 x 1024 cells with dx = 1/128 on (-2, 6) x (-4, 4), 5
nested levels, dt = 1e-3, 402 body points, initial vorticity a Gaussian vortex
of circulation 0.01 and width 0.1 at (1.5, 0.3) (the 5-level 1024^2 precision
configuration run in FP64).  A step is one advance!(state, prob, t).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mixed] [--graphs 0|1]
    python bench.py --kernels-only [--kernel-reps R]   (per-kernel timing only)
    python bench.py --workload naca64 [--steps K]       (batched 64-instance NACA workload)
    python bench.py --workload configs [--config NAME] [--no-cpu] [--graphs 0|1]
                                                        (every configuration, CPU oracle beside it)

Prints one JSON line.  Timed region: K consecutive steps enqueued by
ibpm_advance_n with the state resident in HBM, bracketed by device
synchronisation (the call returns after its stream has drained).  `value` is
the FP64 drop-in's rate; the `mixed` object gives the mixed-precision form
(MFMA sine-matrix transforms) on the same steps from the same start, with its
CD/CL difference from the FP64 run and the MFMA rate of its transforms.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in ("ibpm.jl_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, p))

import numpy as np  # noqa: E402

METRIC = "timesteps/s, 5-level 1024² grid: FP64 vs mixed @CD err≤1e-6; % HBM peak"
HBM_PEAK_GBS = 8000.0
# launch-shape options of the autotuner (ibpm_set_option ids) by name
TUNED_NAMES = {0: "ytile_batch", 1: "ytile_single", 4: "xwaves_batch", 7: "pipe", 8: "ysplit"}


def algorithmic_bytes_per_step(nx, ny, mg, nb):
    """8 [L (4Q + 24N) + 2.5 (L-1) N + 8N + (2 nb)^2] bytes: every operator of
    advance! reads its inputs and writes its output once (static body)."""
    N = (nx - 1) * (ny - 1)
    Q = ny * (nx + 1) + nx * (ny + 1)
    L = mg
    return 8.0 * (L * (4 * Q + 24 * N) + 2.5 * (L - 1) * N + 8 * N + (2 * nb) ** 2)


def measured_traffic(kernel, tile, nlev):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by the copy-kernel
    calibration): the value on its line `traffic <kernel> <tile> <nlev> <bytes>`,
    or None when no summary covers this launch shape."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_cylinder_1024x1024x5.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == "traffic" and f[1] == kernel and f[2] == str(tile) and f[3] == str(nlev):
            return float(f[4])
    return None


def ysingle(tuned, bc):
    """Name of the one-level y-pass launch shape the autotuner chose: tile,
    boundary terms, and two wavefronts per column (split) or one."""
    return (f"k_ypass<{tuned['ytile_single']}{', BC' if bc else ''}"
            f"{', split' if tuned.get('ysplit') else ''}>")


def time_kernels(prob, tuned, reps):
    """Average device ms per launch of each kernel kind (ibpm_time_kernel:
    HIP events on the problem's stream).  Returns ({name: ms}, name of the
    all-level y-pass)."""
    from ibpm import _capi as C

    yname = f"k_ypass<{tuned['ytile_batch']}> op=A (all levels)"
    kb = {}
    for which, name in ((0, "k_xpass<X_PLAIN> (all levels)"), (1, yname), (2, "k_nonlinear (all levels)"),
                        (4, "k_xpass<X_PLAIN> (level 1)"), (5, f"{ysingle(tuned, False)} op=Dinv (level 1)"),
                        (8, f"{ysingle(tuned, True)} op=Dinv + ψ boundary terms (level 1)"),
                        (6, "k_curl (level 1)"), (7, "k_copy (nG*mg doubles, counter calibration)")):
        v = ctypes.c_double(0.0)
        prob.ctx.check(C.lib().ibpm_time_kernel(prob.h, which, reps, ctypes.byref(v)))
        kb[name] = v.value
    return kb, yname


def bench_naca64(args):
    """Batched workload: 64 independent NACA 0012 simulations of uneven grids
    (tests/golden/naca64.json), impulsively started.  value = total
    timesteps/s over all instances with the instances of each grid shape in a
    group (ibpm_group_*: every pass of the step one launch over the group's
    instances; the groups on their own streams).  Beside it the same steps
    with every problem on its own stream (ibpm_advance_async/_wait, graph
    replay) and one problem after another."""
    import collections

    import ibpm
    from ibpm.autotune import autotune
    from cases import naca64_cases

    cs = naca64_cases()
    t0 = time.time()
    probs, states = [], []
    from ibpm import _capi as C

    for c in cs:
        p = ibpm.IBProblem(c["grid"], c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
        autotune(p, reps=5)
        p.set_option(C.OPT_GRAPHS, args.graphs)
        probs.append(p)
        states.append(ibpm.IBState(p))
    setup_s = time.time() - t0
    print(f"# setup of {len(cs)} problems (B assembled on the GPU) {setup_s:.1f}s", file=sys.stderr, flush=True)
    K, W = args.steps, args.warmup
    ts = lambda c, k0, n: (k0 + np.arange(n)) * c["dt"]  # noqa: E731
    ibpm.advance_batch_(states, probs, [ts(c, 0, W) for c in cs])
    # one after another
    t1 = time.perf_counter()
    for st, p, c in zip(states, probs, cs):
        ibpm.advance_n_(st, p, ts(c, W, K))
    el_seq = time.perf_counter() - t1
    # every problem on its own stream
    t2 = time.perf_counter()
    res = ibpm.advance_batch_(states, probs, [ts(c, W + K, K) for c in cs])
    el_str = time.perf_counter() - t2
    if any(r[2] != K for r in res):
        raise RuntimeError("non-finite forces in the batch")
    # groups of the instances of each grid shape, each on its own stream
    shapes = collections.defaultdict(list)
    # one group per grid shape: groups could mix level counts (each walks its
    # deepest member's level chain), but the shorter chains of the 9 shape
    # groups overlapping on their streams measured faster than 3 mixed groups
    for k, c in enumerate(cs):
        shapes[(c["grid"].nx, c["grid"].ny, c["grid"].mg)].append(k)
    members = list(shapes.values())
    groups = [ibpm.Group([probs[k] for k in ks], [states[k] for k in ks]) for ks in members]
    from ibpm.autotune import autotune_group

    for grp in groups:
        grp.set_option(C.OPT_GRAPHS, args.graphs)
        autotune_group(grp)  # the all-level launch shapes of the whole group (its timed steps warm it up)
    for grp in groups:
        grp.advance_async(W)
    for grp in groups:
        grp.wait()
    t3 = time.perf_counter()
    for grp in groups:
        grp.advance_async(K)
    gres = [grp.wait() for grp in groups]
    el_groups = time.perf_counter() - t3
    if any(r[2] != K for g in gres for r in g):
        raise RuntimeError("non-finite forces in a group")
    # the 9 groups in one graph per step (a branch per group), replayed K times
    # (with --graphs 0, for a profiler: the groups on their streams, launched
    # directly, are the value)
    if args.graphs:
        ibpm.advance_groups_(groups, W)
        t4 = time.perf_counter()
        jres = ibpm.advance_groups_(groups, K)
        el = time.perf_counter() - t4
    else:
        jres, el = gres, el_groups
    if any(r[2] != K for g in jres for r in g):
        raise RuntimeError("non-finite forces in the joint step")
    final = {}
    for ks, gr in zip(members, jres):
        for k, (cd, cl, _) in zip(ks, gr):
            final[k] = (float(cd[-1, 0]), float(cl[-1, 0]))
    for grp in groups:
        grp.close()
    total = len(cs) * K
    byts = sum(algorithmic_bytes_per_step(c["grid"].nx, c["grid"].ny, c["grid"].mg, len(c["bodies"][0].ds))
               for c in cs) * K
    out = {
        "metric": "timesteps/s summed over 64 NACA 0012 instances (400x200 / 600x300 / 800x400, 3-5 levels)",
        "value": round(total / el, 1), "unit": "timesteps/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(1e3 * el / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "64 independent NACA 0012 simulations, Re 100-1000, alpha 0-20 deg, dt = 0.25 dx, "
                               "impulsive start; the instances of each grid shape in one group (every pass one "
                               "launch over the group), one step of all groups one HIP graph with a branch per "
                               "group (ibpm_groups_advance_async)",
                   "instances": len(cs), "groups": len(groups), "setup_s": round(setup_s, 1)},
        "groups_on_streams": {"value": round(total / el_groups, 1), "unit": "timesteps/s",
                              "how": "the same groups, each on its own stream, its steps replayed from its own graphs"},
        "streams": {"value": round(total / el_str, 1), "unit": "timesteps/s",
                    "how": "every problem on its own stream (ibpm_advance_async/_wait), graph-replayed steps"},
        "one_by_one": {"value": round(total / el_seq, 1), "unit": "timesteps/s"},
        "step_hbm": {"algorithmic_bytes": byts, "achieved_GBs": round(byts / el / 1e9, 1),
                     "frac_of_peak": round(byts / el / 1e9 / HBM_PEAK_GBS, 4)},
        "per_instance": [{"grid": [c["grid"].nx, c["grid"].ny, c["grid"].mg], "Re": round(c["Re"], 3),
                          "alpha_deg": round(math.degrees(c["freestream"]["inclination"](0.0)), 4),
                          "CD": final[k][0], "CL": final[k][1], "steps_from_start": 3 * W + 4 * K}
                         for k, c in enumerate(cs)],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_configs(args):
    """One line per configuration: GPU FP64 timesteps/s (state resident, one
    ibpm_advance_n call), the same rate as a fraction of 8 TB/s by the bytes
    formula, and the C oracle's timesteps/s on 1 and on 16 host threads
    (a few seconds of steps each, after one untimed step).  --config keeps
    one configuration, --no-cpu skips the oracle legs, --graphs 0 launches
    every kernel from the host (for a profiler)."""
    import ibpm
    import oracle as O
    from ibpm import _capi as C
    from ibpm.autotune import autotune
    from cases import boundary_of, case, gaussian_vortex_gamma, oracle_bodies, oracle_steps

    runs = [("cyl_re40_200x200x3", "cylinder Re=40, 200x200, 3 levels", 600, False),
            ("cyl_re100_800x400x5", "cylinder Re=100, 800x400, 5 levels, Gaussian vortex", 300, True),
            ("cyl_re200_1024x1024x5", "cylinder Re=200, 1024x1024, 5 levels, Gaussian vortex", 300, True),
            ("plate_pitch_512x256x4", "MovingGrid pitching plate Re=200, 512x256, 4 levels", 300, False),
            ("eldredge_lab", "MotionFunction pitching plate Re=100, 400x200, 5 levels (lab frame)", 200, False)]
    if args.config:
        runs = [r for r in runs if r[0] == args.config]
        if not runs:
            raise SystemExit(f"unknown configuration {args.config}")
    for name, desc, steps, vortex in runs:
        c = case(name)
        g = c["grid"]
        nb = sum(len(b.ds) for b in c["bodies"])
        motion = 2 if isinstance(c["bodies"][0].motion, ibpm.MotionFunction) else (1 if "motion" in c else 0)
        prob = ibpm.IBProblem(g, c["bodies"], c["dt"], c["Re"], freestream=c["freestream"])
        autotune(prob, reps=5)
        prob.set_option(C.OPT_GRAPHS, args.graphs)
        st = ibpm.IBState(prob)
        if vortex:
            st.Γ = gaussian_vortex_gamma(g)
            ibpm.vort2flux_(st, prob)
        w = 20
        ibpm.advance_n_(st, prob, np.arange(w) * c["dt"])
        ts = (w + np.arange(steps)) * c["dt"]
        t0 = time.perf_counter()
        _, _, done = ibpm.advance_n_(st, prob, ts)
        el = time.perf_counter() - t0
        sps = done / el
        byts = algorithmic_bytes_per_step(g.nx, g.ny, g.mg, nb) + (8.0 * g.mg * g.nq if motion == 1 else 0.0)
        cpu = {1: None, 16: None}
        for threads in (() if args.no_cpu else (1, 16)):
            O.set_threads(threads)
            op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], c["dt"], oracle_bodies(c), motion=motion,
                                 binv=None if motion == 2 else prob.Binv)
            S = op.new_state(oracle_steps(c, [0.0])[0])
            if vortex:
                S.G[:] = gaussian_vortex_gamma(g)
            rows = oracle_steps(c, np.arange(12) * c["dt"])
            op.advance(S, rows[0])
            k, tc = 0, time.perf_counter()
            while k < 10 and (time.perf_counter() - tc) < 3.0:
                k += 1
                op.advance(S, rows[k])
            cpu[threads] = k / (time.perf_counter() - tc)
        print(json.dumps({"workload": name, "desc": desc, "grid": [g.nx, g.ny], "levels": g.mg, "body_points": nb,
                          "gpu_timesteps_per_s": round(sps, 1), "steps": done,
                          "hbm_frac_by_bytes_formula": round(byts * sps / 1e9 / HBM_PEAK_GBS, 4),
                          "cpu_oracle_1_thread": cpu[1] and round(cpu[1], 3),
                          "cpu_oracle_16_threads": cpu[16] and round(cpu[16], 3)}),
              flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--graphs", type=int, default=1,
                    help="0: every step launches its kernels (no HIP-graph replay), e.g. under a profiler")
    ap.add_argument("--cpu-steps", type=int, default=12,
                    help="oracle steps in the CPU baseline sample (about 1 s each on one core)")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--workload", choices=("cylinder1024", "naca64", "configs"), default="cylinder1024",
                    help="naca64: the batched 64-instance NACA workload (secondary line); configs: one line "
                         "per configuration with the CPU oracle beside it")
    ap.add_argument("--config", default=None,
                    help="with --workload configs: only the named configuration (e.g. cyl_re100_800x400x5)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only autotune and time each kernel kind (for counter profiling); prints kernel_ms")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (counter passes over the steps then launch the same "
                         "kernels in every pass)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "naca64":
        return bench_naca64(args)
    if args.workload == "configs":
        return bench_configs(args)

    import ibpm
    from ibpm import _capi as C
    from cases import case, gaussian_vortex_gamma

    c = case("cyl_re200_1024x1024x5")
    g = c["grid"]
    nb = len(c["bodies"][0].ds)
    dt = c["dt"]
    from ibpm.autotune import autotune

    if args.kernels_only:
        # B's values change no kernel's memory traffic: take it as given (identity)
        prob = ibpm.IBProblem(g, c["bodies"], dt, c["Re"], freestream=c["freestream"], Binv=np.eye(2 * nb))
        state = ibpm.IBState(prob)
        state.Γ = gaussian_vortex_gamma(g)
        ibpm.vort2flux_(state, prob)
        tuned = {TUNED_NAMES[k]: v[0] for k, v in autotune(prob).items()}
        kb, _ = time_kernels(prob, tuned, args.kernel_reps)
        # the mixed-precision sine-matrix transforms (MFMA), both arithmetics
        for prec in ("bf16x3", "fp32"):
            prob.set_precision(prec)
            for which, name in ((9, "x-transform (all levels)"), (10, "y-transform + eigenvalue scaling (all levels)"),
                                (11, "x-transform (level 1)")):
                v = ctypes.c_double(0.0)
                prob.ctx.check(C.lib().ibpm_time_kernel(prob.h, which, args.kernel_reps, ctypes.byref(v)))
                kb[f"k_mx_gemm {prec} {name}"] = v.value
        prob.set_precision("fp64")
        print(json.dumps({"kernels_only": True, "tiles": tuned, "kernel_ms": {k: round(v, 5) for k, v in kb.items()}}),
              flush=True)
        return
    t0 = time.time()
    prob = ibpm.IBProblem(g, c["bodies"], dt, c["Re"], freestream=c["freestream"])
    setup_s = time.time() - t0
    prob.set_option(C.OPT_GRAPHS, args.graphs)
    state = ibpm.IBState(prob)
    state.Γ = gaussian_vortex_gamma(g)
    ibpm.vort2flux_(state, prob)

    if args.no_autotune:
        tuned = {name: prob.get_option(k) for k, name in TUNED_NAMES.items()}
    else:
        tuned = {TUNED_NAMES[k]: v[0] for k, v in autotune(prob).items()}
    print(f"# setup {setup_s:.2f}s (B assembly on GPU {prob.info.setup_seconds:.2f}s, "
          f"|B-B'|/|B| = {prob.info.b_asymmetry:.2e}); tiles {tuned}; warmup {args.warmup} steps",
          file=sys.stderr, flush=True)
    ts_w = np.arange(args.warmup) * dt
    ibpm.advance_n_(state, prob, ts_w)
    ts = (args.warmup + np.arange(args.steps)) * dt
    t1 = time.perf_counter()
    cd, cl, done = ibpm.advance_n_(state, prob, ts)
    el = time.perf_counter() - t1
    if done != args.steps:
        raise RuntimeError(f"non-finite forces after {done} steps")
    # the roofline kernels are timed inside further, untimed steps (their event
    # pairs would slow the timed ones): the one-level y-passes with boundary
    # terms of the closing vort2flux! chain (profile kind 2), then every
    # all-level y-pass (kind 1), bracketed by HIP events on the problem's stream
    n_extra = min(args.steps, 100)
    t_next = args.warmup + args.steps
    prof = {}
    for kind in (2, 1):
        prob.ctx.check(C.lib().ibpm_set_option(prob.h, C.OPT_PROFILE, kind), "ibpm_set_option")
        ibpm.advance_n_(state, prob, (t_next + np.arange(n_extra)) * dt)
        t_next += n_extra
        ms, n = ctypes.c_double(0.0), ctypes.c_int64(0)
        prob.ctx.check(C.lib().ibpm_profile_stats(prob.h, ctypes.byref(ms), ctypes.byref(n)), "ibpm_profile_stats")
        prof[kind] = (ms.value, n.value)
    y_ms, y_n = prof[2]
    a_ms, a_n = prof[1]
    prob.ctx.check(C.lib().ibpm_set_option(prob.h, C.OPT_PROFILE, 0), "ibpm_set_option")
    sps = args.steps / el
    print(f"# timed {args.steps} steps in {el:.3f}s: {sps:.1f} steps/s; CD={cd[-1, 0]:.6f} CL={cl[-1, 0]:.3e}",
          file=sys.stderr, flush=True)

    # dominant-kernel roofline, HIP events on the problem's stream
    L = g.mg
    N = g.nΓ
    kb, yname = time_kernels(prob, tuned, args.kernel_reps)
    # algorithmic bytes per launch: read N + write N doubles per level for a DST
    # pass; the boundary-term pass also reads the two coarse columns (one
    # pair per row) and the two coarse rows (nx/2 + 1 samples each).  Of the
    # L - 1 timed launches, level 1's reads a second field (its input is
    # 2 nx (Ainv's y-pass output - the projection's)): N more doubles, shown
    # here as the average over the timed launches
    bc_bytes = 2.0 * 8.0 * N + 8.0 * (2 * (g.ny - 1) + 2 * (g.nx // 2 + 1)) + 8.0 * N / (L - 1)
    ypass_ms = y_ms
    roof = {"bound": "hbm",
            "kernel": f"{ysingle(tuned, True)} (one-level Δinv y-pass with the ψ boundary terms; "
                      f"{2 * (L - 1)} per step, the vort2flux! chains' levels below the coarsest, the largest "
                      f"share of a step's kernel time; timed: the {L - 1} of the step's closing vort2flux!, which "
                      f"read and write the whole level (the trial chain's store only the rows its next pass reads))",
            "achieved": round(bc_bytes / (ypass_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bc_bytes / (ypass_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("k_ypass_bc_split" if tuned.get("ysplit") else "k_ypass_bc",
                                        tuned["ytile_single"], 1),
            "bytes_per_launch": bc_bytes, "avg_launch_ms": round(ypass_ms, 5), "launches_timed": y_n}
    dst_bytes = 2.0 * 8.0 * N * L
    roof_all = {"bound": "hbm",
                "kernel": f"k_ypass<{tuned['ytile_batch']}> (fused DST-I y-pass, all 5 levels; Ainv of every step)",
                "achieved": round(dst_bytes / (a_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dst_bytes / (a_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("k_ypass", tuned["ytile_batch"], L),
                "bytes_per_launch": dst_bytes, "avg_launch_ms": round(a_ms, 5), "launches_timed": a_n}
    step_bytes = algorithmic_bytes_per_step(g.nx, g.ny, g.mg, nb)

    mixed = None if args.no_mixed else bench_mixed(prob, c, args, cd, cl)

    cpu = None
    if not args.no_cpu:
        import oracle as O
        from cases import boundary_of, oracle_bodies, sc7

        def oracle_rate(threads, nsteps):
            O.set_threads(threads)
            op = O.OracleProblem(g.h, boundary_of(g), g.mg, c["Re"], dt, oracle_bodies(c), binv=prob.Binv)
            S = op.new_state(sc7(c, 0.0))
            S.G[:] = gaussian_vortex_gamma(g)
            op.run(S, sc7(c, 0.0)[None, :])  # first step outside the timing
            tc = time.perf_counter()
            op.run(S, np.tile(sc7(c, 0.0), (nsteps, 1)))
            return nsteps, time.perf_counter() - tc

        # all the host cores this job may use (the box's share, OMP_NUM_THREADS), then one core
        nth = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        n_all, e_all = oracle_rate(nth, 4 * args.cpu_steps)
        n_one, e_one = oracle_rate(1, args.cpu_steps)
        cpu = {"value": round(n_all / e_all, 4), "unit": "timesteps/s", "cores": nth, "kind": "port",
               "sample": f"{n_all} advance! steps of the same 1024x1024 5-level case (after 1 untimed), C oracle, "
                         f"{nth} OpenMP threads, {e_all:.1f}s",
               "one_core": {"value": round(n_one / e_one, 4), "cores": 1,
                            "sample": f"{n_one} steps on 1 thread, {e_one:.1f}s"}}

    out = {
        "metric": METRIC, "value": round(sps, 2), "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "FP64 advance!, cylinder Re=200, 1024x1024 cells (dx=1/128 on [-2,6]x[-4,4]), "
                               "5 levels, dt=1e-3, 402 body points, Gaussian-vortex start",
                   "grid": [g.nx, g.ny], "levels": g.mg, "body_points": nb, "precision": "fp64",
                   "tiles": tuned},
        "roofline": roof,
        "roofline_all_level": roof_all,
        "step_hbm": {"algorithmic_bytes_per_step": step_bytes,
                     "achieved_GBs": round(step_bytes * sps / 1e9, 1),
                     "frac_of_peak": round(step_bytes * sps / 1e9 / HBM_PEAK_GBS, 4)},
        # isolated repeated launches (operands stay in the 256 MB Infinity Cache): faster than in-step
        "kernel_ms_isolated": {k: round(v, 5) for k, v in kb.items()},
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_mixed(prob, c, args, cd64, cl64):
    """The mixed-precision form on the same problem: from the same start as
    the FP64 run, the same warmup and timed steps in split BF16 and in FP32
    (MFMA sine-matrix transforms, two sweeps per solve); rate, ratio to the
    FP64 rate and max |CD - CD_fp64|, |CL - CL_fp64| over the timed steps
    relative to max |CD_fp64|; and the MFMA rate of the all-level x-transform
    (the transforms' hardware operations per launch / its average duration,
    HIP events on the problem's stream)."""
    import ibpm
    from ibpm import _capi as C
    from cases import gaussian_vortex_gamma

    g = c["grid"]
    dt = c["dt"]
    P1 = (g.nx - 1 + 63) // 64 * 64
    P2 = (g.ny - 1 + 63) // 64 * 64
    res = {}
    fp64_rate = None
    for prec, name, peak in (("bf16x3", "split BF16 (hi*hi + hi*lo + lo*hi), FP32 accumulation", 2500.0),
                             ("fp32", "FP32", 157.3)):
        prob.set_precision(prec, 2)
        st = ibpm.IBState(prob)
        st.Γ = gaussian_vortex_gamma(g)
        ibpm.vort2flux_(st, prob)
        ibpm.advance_n_(st, prob, np.arange(args.warmup) * dt)
        ts = (args.warmup + np.arange(args.steps)) * dt
        t1 = time.perf_counter()
        cd, cl, done = ibpm.advance_n_(st, prob, ts)
        el = time.perf_counter() - t1
        scale = float(np.max(np.abs(cd64)))
        v = ctypes.c_double(0.0)
        prob.ctx.check(C.lib().ibpm_time_kernel(prob.h, 9, args.kernel_reps, ctypes.byref(v)))
        ops = 2.0 * P1 * P1 * P2 * g.mg * (3 if prec == "bf16x3" else 1)
        tf = ops / (v.value * 1e-3) / 1e12
        res[prec] = {"value": round(done / el, 2), "unit": "timesteps/s", "arithmetic": name, "sweeps": 2,
                     "cd_err_vs_fp64": float(np.max(np.abs(cd - cd64)) / scale),
                     "cl_err_vs_fp64": float(np.max(np.abs(cl - cl64)) / scale),
                     "mfma": {"kernel": "k_mx_gemm x-transform, all levels", "achieved": round(tf, 1),
                              "peak": peak, "unit": "TFLOP/s", "frac": round(tf / peak, 4),
                              "avg_launch_ms": round(v.value, 5)}}
    prob.set_precision("fp64")
    return res


if __name__ == "__main__":
    main()
