/*
 * ibpm_hip.h -- C ABI of the MI355X (gfx950) implementation of IBPM.jl's
 * `advance!(state, prob, t)` hot path.
 *
 * The reference has no FFI for this path: its boundary is the Julia method
 *   advance!(state::IBState, prob::AbstractIBProblem, t::Float64)
 * (src/timestepping/timestepping.jl:35-67), with state and problem built by
 *   IBProblem(grid, bodies, dt, Re; freestream)  (src/pre-processing/problem-types.jl:23-41)
 *   IBState(prob)                                (src/pre-processing/state-types.jl:25-47)
 * The functions below are what a `ccall` layer binding that method would
 * call.  Every array crossing this ABI uses the reference's exact memory
 * layout (Julia column-major, one column per grid level):
 *   q, q0           : nq x mg     nq = ny(nx+1) + nx(ny+1); x-fluxes first,
 *                                  as (nx+1) x ny, then y-fluxes as nx x (ny+1)
 *   Γ, ψ, nonlin[k] : nΓ x mg     nΓ = (nx-1)(ny-1)
 *   F̃b              : 2 nb_total  [fx(1..nb_total); fy(1..nb_total)]
 *   fb              : for each body j its (nb_j x 2) block, bodies in order
 *   CD, CL          : one value per body
 * (src/fluid-domain/fluid-domain.jl:44-73, src/pre-processing/state-types.jl:7-47).
 *
 * Conventions: every function returns 0 on success or a negative IBPM_E_*
 * code; the message of the last failure on a context is returned by
 * ibpm_last_error().  Host pointers are borrowed for the duration of the call
 * only.  The library owns all device memory.  Calls on one problem must be
 * serialised (the reference's `advance!` is likewise non-reentrant: it works
 * in prob.model.work, src/pre-processing/models.jl:79-106); different problems
 * may be driven from different host threads.
 */
#ifndef IBPM_HIP_H
#define IBPM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibpm_ctx ibpm_ctx;
typedef struct ibpm_prob ibpm_prob;
typedef struct ibpm_group ibpm_group;

/* status codes */
#define IBPM_OK 0
#define IBPM_E_ARG -1        /* invalid argument / size mismatch */
#define IBPM_E_HIP -2        /* HIP runtime failure */
#define IBPM_E_UNSUPPORTED -3/* motion type or grid shape the library does not run */
#define IBPM_E_SINGULAR -4   /* Schur matrix B singular during setup */
#define IBPM_E_NONFINITE -5  /* non-finite body force after a step (ibpm.jl:95-97 test) */

/* motion kinds: MotionType(bodies) (src/structure-domain/move-body-utils.jl:15-26) */
#define IBPM_MOTION_STATIC 0
#define IBPM_MOTION_MOVING_GRID 1
#define IBPM_MOTION_FUNCTION 2   /* lab-frame moving body (motion-types.jl:14-17): E is rebuilt and the
                                    Schur system solved by CG every step */

/* state fields for upload/download (IBState, src/pre-processing/state-types.jl:7-20) */
#define IBPM_FIELD_Q 0
#define IBPM_FIELD_Q0 1
#define IBPM_FIELD_GAMMA 2
#define IBPM_FIELD_PSI 3
#define IBPM_FIELD_NONLIN1 4
#define IBPM_FIELD_NONLIN2 5
#define IBPM_FIELD_FB_TILDE 6   /* F̃b */
#define IBPM_FIELD_FB 7         /* fb, all bodies stacked */
#define IBPM_FIELD_CD 8
#define IBPM_FIELD_CL 9

/* Grid + scheme description: make_grid(h, boundary; mg)
 * (src/fluid-domain/fluid-domain.jl:44-73), Re and dt of IBProblem
 * (problem-types.jl:29-33), AB2 coefficients (models.jl:70-72). */
typedef struct {
    int32_t nx, ny, mg;
    double h;            /* grid.h = len/nx */
    double offx, offy;   /* grid.offx, grid.offy */
    double Re, dt;
    double beta[2];      /* AdamsBashforth β; AB2 = {1.5, -0.5} */
} ibpm_grid_desc;

/* Bodies: vcat of RigidBody.xb / .ds (src/structure-domain/body-types.jl:10-16).
 * For IBPM_MOTION_FUNCTION, xb are also the reference points x0 that
 * move_body! transforms (x0 = copy(xb), sample-bodies.jl:32). */
typedef struct {
    int32_t nbodies;
    const int32_t* nb;   /* points per body */
    const double* xb;    /* nb_total x 2, column-major (all x, then all y) */
    const double* ds;    /* nb_total */
    int32_t motion;      /* IBPM_MOTION_* */
} ibpm_body_desc;

/* Closures of the problem evaluated by the caller at the step time t
 * (freestream: src/pre-processing/problem-types.jl:33; MovingGrid U,V,θ,θ̇:
 * src/structure-domain/motion-types.jl:42-50). */
typedef struct {
    double Ux, Uy, inclination;      /* freestream.Ux(t), .Uy(t), .inclination(t) */
    double U, V, theta, theta_dot;   /* motion.U(t), .V(t), .θ(t), .θ̇(t) (MovingGrid only) */
    /* MotionFunction only: 6 values per body, {x, y, θ, ẋ, ẏ, θ̇} =
     * motion.xc[1:3](t), motion.uc[1:3](t) (move-body-utils.jl:73-97); read
     * during the call that takes these scalars.  NULL otherwise. */
    const double* body_kin;
} ibpm_step_scalars;

/* Sizes and setup products of a problem (for host mirrors and tests). */
typedef struct {
    int32_t nx, ny, mg, nbodies, nb_total, nbc;
    int64_t nG, nq;
    double h, offx, offy;
    double b_asymmetry;      /* ||B - B'||_F / ||B||_F of the assembled Schur matrix (0 if Binv given) */
    double setup_seconds;
} ibpm_problem_info;

/* Context on one device.  Replaces nothing in the reference (it is the HIP
 * device binding the Julia module would hold). */
int ibpm_ctx_create(int device, ibpm_ctx** out);
int ibpm_ctx_destroy(ibpm_ctx* ctx);
const char* ibpm_last_error(const ibpm_ctx* ctx);

/* IBProblem(grid, bodies, dt, Re; freestream) -- problem-types.jl:23-41,
 * including IBModel/IBMatrices (models.jl:33-52,124-158: Λ, the DST plans,
 * E = setup_reg, interface-coupling.jl:72-96) and get_AB (laplacian.jl:83-89).
 * binv: if non-NULL, the (2 nb_total)^2 column-major prob.Binv to use; if
 * NULL, B is assembled on the GPU column by column exactly as get_Binv
 * (laplacian.jl:98-119, B_times! :143-169) and inverted in FP64 on the host
 * (LU with partial pivoting, as inv() does). */
int ibpm_problem_create(ibpm_ctx* ctx, const ibpm_grid_desc* grid, const ibpm_body_desc* bodies,
                        const double* binv, ibpm_prob** out);
/* For IBPM_MOTION_FUNCTION bodies there is no matrix: prob.Binv is the CG map
 * of get_Binv (laplacian.jl:123-141), so `binv` must be NULL and
 * ibpm_problem_get_binv returns IBPM_E_UNSUPPORTED. */
int ibpm_problem_destroy(ibpm_prob* prob);
/* LinearizedIBProblem(base_state, base_prob, dt) -- problem-types.jl:47-87:
 * turns `prob` (built with the base problem's grid, bodies, Re and this dt)
 * into the problem linearized about the base state's q, q0 and Γ (nq x mg,
 * nq x mg, nG x mg, column-major).  QB = avg_flux(base_state, base_prob) per
 * level and ΓB = base Γ are kept on the device; nonlinear! then forms
 * Q x ΓB + QB x Γ (nonlinear.jl:45-65,144-163) and base_flux! sets q0 = 0
 * (timestepping.jl:238-250, also applied now).  Static bodies only. */
int ibpm_problem_linearize(ibpm_prob* prob, const double* qB, const double* q0B, const double* GammaB);
int ibpm_problem_get_info(const ibpm_prob* prob, ibpm_problem_info* info);
/* prob.Binv (column-major), body_idx (nb_total x 2, column-major, Julia
 * 1-based values as computed by setup_reg, interface-coupling.jl:85-87). */
int ibpm_problem_get_binv(ibpm_prob* prob, double* out, size_t n);
int ibpm_problem_get_body_idx(ibpm_prob* prob, int32_t* out, size_t n);

/* IBState fields <-> device (state-types.jl:7-20).  n = element count. */
int ibpm_state_upload(ibpm_prob* prob, int field, const double* host, size_t n);
int ibpm_state_download(ibpm_prob* prob, int field, double* host, size_t n);
/* IBState(prob): zero every field, then base_flux!(state, prob, 0)
 * (state-types.jl:25-47; timestepping.jl:252-270 / 273-318). */
int ibpm_state_init(ibpm_prob* prob, const ibpm_step_scalars* at_t0);
/* base_flux!(state, prob, t) (timestepping.jl:230-318). */
int ibpm_base_flux(ibpm_prob* prob, const ibpm_step_scalars* at_t);
/* vort2flux!(state.ψ, state.q, state.Γ, model, mg) on the device state
 * (curl.jl:133-165): makes ψ and q consistent with an uploaded Γ. */
int ibpm_vort2flux(ibpm_prob* prob);

/* advance!(state, prob, t) -- timestepping.jl:35-67.  `s` holds the problem's
 * closures evaluated at t (the time at the start of the step).  Returns when
 * the step is done; IBPM_E_NONFINITE if its CL came out non-finite (the state
 * is still updated). */
int ibpm_advance(ibpm_prob* prob, double t, const ibpm_step_scalars* s);
/* nsteps calls of advance! with per-step scalars s[k] (times t[k]) without a
 * host round trip in between: the time loop of run_sim! (ibpm.jl:93-105).
 * cd_hist / cl_hist (nsteps x nbodies, row k = step k) may be NULL.  Stops at
 * the first step whose CL is non-finite (ibpm.jl:95-97) and returns
 * IBPM_E_NONFINITE with *steps_done set. */
int ibpm_advance_n(ibpm_prob* prob, int nsteps, const double* t, const ibpm_step_scalars* s,
                   double* cd_hist, double* cl_hist, int* steps_done);
/* ibpm_advance_n split in two, so that several problems (each with its own
 * stream) can step concurrently: ibpm_advance_async enqueues nsteps steps and
 * returns at once; ibpm_advance_wait drains the problem's stream and returns
 * the CD/CL histories of every step enqueued since the last wait (same
 * layout and non-finite rule as ibpm_advance_n).  No counterpart in the
 * reference: it is the batched form of run_sim!'s loop (ibpm.jl:93-105). */
int ibpm_advance_async(ibpm_prob* prob, int nsteps, const ibpm_step_scalars* s);
int ibpm_advance_wait(ibpm_prob* prob, double* cd_hist, double* cl_hist, int* steps_done);
/* MotionFunction problems step synchronously inside these calls (the CG
 * solve reads its residual every iteration): ibpm_advance_async then returns
 * after its steps are done, and ibpm_advance_wait only collects them.  A step
 * fails with IBPM_E_ARG, leaving the state and the bodies as they were
 * before that step, when a body point would come within 6 cells of the
 * domain edge (where the reference's E window raises a BoundsError). */

/* state.xb and RigidBody.ub after the last update_bodies! (timestepping.jl:386-398):
 * nb_total x 2 column-major each (n = 2 nb_total); either may be NULL.  For
 * Static and MovingGrid bodies, xb are the points given at creation, ub = 0. */
int ibpm_get_body_state(ibpm_prob* prob, double* xb, double* ub, size_t n);
/* MotionFunction: products B u used and |r| reached by the last step's CG
 * solve, and its tolerance 1e-12 |b| (cg!, laplacian.jl:137-138).  No
 * counterpart in the reference (cg! is called without log = true there). */
int ibpm_cg_stats(ibpm_prob* prob, int32_t* iterations, double* residual, double* tolerance);

/* Groups: n problems with static bodies on the same grid (nx, ny, mg, h,
 * offsets; dt, Re, freestream and bodies may differ) stepped together, each
 * pass of advance! one launch over all instances.  No counterpart in the
 * reference: it is the batched form of many run_sim! loops (ibpm.jl:93-105)
 * over independent simulations.  While grouped, a member's state lives in the
 * group's arrays: it is read with ibpm_state_download / ibpm_get_forces (after
 * ibpm_group_wait), and the member's own stepping, upload, state_init,
 * base_flux, vort2flux, linearize and destroy calls return IBPM_E_ARG.
 * ibpm_group_destroy copies every member's state back and releases them.
 * Every instance takes the same steps: the results are those of stepping each
 * member alone. */
int ibpm_group_create(ibpm_ctx* ctx, int n, ibpm_prob* const* members, ibpm_group** out);
int ibpm_group_destroy(ibpm_group* group);
/* IBPM_OPT_YTILE_*, IBPM_OPT_XWAVES_BATCH and IBPM_OPT_GRAPHS (capture of the
 * group's steps in HIP graphs, on by default) for the group's launches */
int ibpm_group_set_option(ibpm_group* group, int option, int value);
/* Enqueue nsteps steps of every member (static bodies: no per-step scalars). */
int ibpm_group_advance_async(ibpm_group* group, int nsteps);
/* Several groups (for instance one per grid shape) stepped together: one step
 * of every group is captured in a single HIP graph, a branch per group (each
 * group's launches on its own stream, forked from and joined into the first
 * group's stream), and the graph is replayed nsteps times.  Collect each
 * group's histories with ibpm_group_wait as after ibpm_group_advance_async.
 * No counterpart in the reference (many run_sim! loops, ibpm.jl:93-105). */
int ibpm_groups_advance_async(int ngroups, ibpm_group* const* groups, int nsteps);
/* Drain; cd_hist / cl_hist (may be NULL) receive, member after member, each
 * member's (steps x nbodies) history of the steps since the last wait;
 * steps_done (n entries, may be NULL) as in ibpm_advance_wait per member;
 * IBPM_E_NONFINITE if any member's CL became non-finite. */
int ibpm_group_wait(ibpm_group* group, double* cd_hist, double* cl_hist, int32_t* steps_done);

/* Tangent-linear and adjoint of advance! -- the discrete derivative of
 *   J = sum_k (wcd[k] CD_k + wcl[k] CL_k)
 * over nsteps steps from the current state (timestepping.jl:35-67 with the
 * MovingGrid base flux :273-318), for one MovingGrid body and the FP64 path.
 * A step's scalars are its 7 doubles {Ux, Uy, inclination, U, V, θ, θ̇} in the
 * order of ibpm_step_scalars; ds / sbar hold nsteps rows of them.  No
 * counterpart in the reference: these are the pieces a ChainRulesCore rrule
 * of the run would call.  Both leave the problem's state as they found it.
 *
 * ibpm_tangent_n: J and dJ = the directional derivative for perturbations ds
 * of the scalars and dGamma0 (nG x mg, may be NULL) of the initial Γ,
 * computed by tangent-linear steps alongside the forward ones.
 * ibpm_adjoint_n: J, sbar = dJ/ds for every step, and (if non-NULL)
 * Gamma0bar = dJ/dΓ of the initial state, by adjoint steps in reverse with
 * binomial checkpointing of (Γ, nonlin[2]) in nsnap snapshots (nsnap >= 1;
 * every reverse step recomputes its input state from the nearest snapshot).
 * forward_steps (may be NULL): forward steps taken, the first run included. */
int ibpm_tangent_n(ibpm_prob* prob, int nsteps, const ibpm_step_scalars* s, const double* ds, const double* dGamma0,
                   const double* wcd, const double* wcl, double* J, double* dJ);
int ibpm_adjoint_n(ibpm_prob* prob, int nsteps, const ibpm_step_scalars* s, const double* wcd, const double* wcl,
                   int nsnap, double* J, double* sbar, double* Gamma0bar, int64_t* forward_steps);

/* state.CD, state.CL (nbodies each) and state.F̃b (2 nb_total); any may be NULL. */
int ibpm_get_forces(ibpm_prob* prob, double* CD, double* CL, double* Fb);
/* compute_cfl(state, prob) = max|q[:,1]| dt / h (ibpm.jl:83-88). */
int ibpm_compute_cfl(ibpm_prob* prob, double* cfl);

/* --- operator-level entry points (unit parity tests; host arrays in/out) --- */
/* kind: 0 = Δinv (laplacian.jl:24-41), 1 = A[lev] (:73-76), 2 = Ainv[lev] (:61-71) */
int ibpm_op_lapinv(ibpm_prob* prob, int kind, int lev, const double* b, double* x);
/* ibpm_op_lapinv for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  It
 * has no counterpart in the reference.  b and x are DEVICE pointers of
 * n = (nx-1)(ny-1) doubles (x may not alias b); the transform between the
 * passes goes through a fixed scratch array of IBPM_NOALLOC_MAX_POINTS
 * doubles (8 MiB) that each device holds for the life of the process, so the
 * call makes no hipMalloc, hipFree or other allocating call.  The work runs on
 * the problem's stream and is waited for before the call returns.  n above
 * the maximum (or different from the problem's level size) gives IBPM_E_ARG.
 * Calls that overlap on one GPU (from several host threads) race on the
 * shared scratch; the scratch keeps the last call's contents, all of which
 * the next call overwrites before reading. */
#define IBPM_NOALLOC_MAX_POINTS 1048576
int ibpm_op_lapinv_noalloc(ibpm_prob* prob, int kind, int lev, const double* b, double* x, size_t n);
/* 2D RODFT00 of one (nx-1) x (ny-1) array (dst-inversion.jl:10-20) */
int ibpm_op_dst2(ibpm_prob* prob, const double* x, double* y);
/* nonlinear!(out, state, bc(lev+1), lev) on host copies of q, q0, Γ (nonlinear.jl:168-201) */
int ibpm_op_nonlinear(ibpm_prob* prob, int lev, const double* q, const double* q0, const double* G, double* out);
/* vort2flux! on host copies (curl.jl:133-165); G is updated (coarse levels) */
int ibpm_op_vort2flux(ibpm_prob* prob, double* G, double* psi, double* q);
/* E (regT!, interface-coupling.jl:122-133) and E' (reg!, :99-119) */
int ibpm_op_regT(ibpm_prob* prob, const double* q, double* fb);
int ibpm_op_reg(ibpm_prob* prob, const double* fb, double* q);
/* rot! (curl.jl:22-34) */
int ibpm_op_rot(ibpm_prob* prob, const double* q, double* G);
/* MotionFunction: update_bodies! alone (timestepping.jl:386-398): move_body!
 * of every body to kin (6 values per body, as body_kin) and E = setup_reg. */
int ibpm_op_update_bodies(ibpm_prob* prob, const double* kin);

/* --- tuning and timing (bench, autotuner) --- */
/* Launch-shape options of a problem.  No counterpart in the reference (it has
 * no kernels); used by the autotuner and the benchmark. */
#define IBPM_OPT_YTILE_BATCH 0   /* columns per y-pass workgroup, all-level launches: 4, 8 or 12 */
#define IBPM_OPT_YTILE_SINGLE 1  /* columns per y-pass workgroup, one-level launches: 4, 8 or 12 */
#define IBPM_OPT_PROFILE 2       /* 1: time every all-level y-pass inside the steps; 2: time every
                                    one-level y-pass with the ψ boundary terms of the vort2flux! chain
                                    (HIP events on the problem's stream); setting it drains the stream
                                    and clears the sums */
#define IBPM_OPT_GRAPHS 3        /* 1 (default): static-body steps are captured once per buffer role in a
                                    HIP graph and replayed; 0: every step launches its kernels */
#define IBPM_OPT_XWAVES_BATCH 4  /* lines per work-group of the plain x-passes over all levels: 4 (default),
                                    8, 12 or 14 (one LDS copy of the transform tables for more lines) */
#define IBPM_OPT_PRECISION 5     /* arithmetic of the Laplacian inverses in advance!: 0 FP64 (default, the
                                    drop-in); 1 every DST-I direction a dense sine-matrix product on MFMA in
                                    FP32; 2 the same in split BF16 (high and low parts, three products,
                                    FP32 accumulation).  In 1 and 2 the transforms act on FP64 residuals of
                                    the state (increment form) and the eigenvalue scaling, residuals,
                                    updates, Schur solve and force sums stay FP64.  Static and MovingGrid
                                    bodies; ibpm_op_lapinv (kinds 0 and 2) follows the setting. */
#define IBPM_OPT_MX_SWEEPS 6     /* sweeps of each mixed-precision solve: 1, or 2 (default): a second
                                    application of the transforms to the FP64 residual of the first */
#define IBPM_OPT_PIPE 7          /* 1 (default): passes over more tiles than the chip holds at once run as
                                    persistent work-groups that load their next tile while transforming the
                                    current one; 0: one work-group per tile */
#define IBPM_OPT_YSPLIT 8        /* 1 (default): the one-level y-passes over 1023-point columns (ny = 1024,
                                    4-column tiles) transform each column with two wavefronts at once (its
                                    even and its odd outputs); 0: one wavefront per column */
#define IBPM_OPT_XREUSE 9        /* 1: the step's vort2flux! chains x-transform only the lines coarsify!
                                    rewrites and read the others as 2 nx times Ainv's y-pass output (psi
                                    then equals vort2flux!(Gamma) up to rounding); 0: the full x-transform.
                                    Default 1 for grids of nx ny >= 2^19 cells, else 0 */
int ibpm_set_option(ibpm_prob* prob, int option, int value);
int ibpm_get_option(ibpm_prob* prob, int option, int* value);
/* Average device ms and count of the y-pass launches timed inside the steps
 * (the kind IBPM_OPT_PROFILE selects) since that option was last set.  No counterpart in the
 * reference (it has no kernels); the benchmark's roofline reads it. */
int ibpm_profile_stats(ibpm_prob* prob, double* avg_ms, int64_t* count);

/* Average device time in ms of `reps` launches of kernel `which` on the
 * problem's stream, bracketed by HIP events on that stream.  which:
 * 0 = x-pass DST over all levels, 1 = fused y-pass (A operator) over all
 * levels, 2 = nonlinear over all levels, 4 = x-pass on level 1 only,
 * 5 = fused y-pass (Δinv) on level 1 only, 6 = curl on level 1,
 * 7 = plain copy of nΓ*mg doubles (HBM-traffic counter calibration),
 * 8 = Δinv y-pass with the ψ boundary terms (vort2flux!'s apply_bc!) on level 1;
 * with IBPM_OPT_PRECISION set: 9 = sine-matrix x-transform (MFMA) over all
 * levels, 10 = y-transform with the eigenvalue scaling over all levels,
 * 11 = x-transform on level 1. */
int ibpm_time_kernel(ibpm_prob* prob, int which, int reps, double* ms_avg);

#ifdef __cplusplus
}
#endif
#endif /* IBPM_HIP_H */
