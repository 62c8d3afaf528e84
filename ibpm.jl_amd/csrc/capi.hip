// capi.hip -- host side of the library: problem setup, the advance! step
// schedule and the C ABI declared in include/ibpm_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ibpm_hip.h"
#include "adjoint.h"
#include "kernels.h"
#include "mixed.h"

using namespace ibpm;

struct ibpm_ctx {
    int device = 0;
    std::string err;
    // device address of this device's copy of `scratch` (ibpm_op_lapinv_noalloc)
    double* scratch = nullptr;
};

// Scratch of ibpm_op_lapinv_noalloc: the x-transformed level between the
// passes of one Laplacian-inverse application.  Each device holds its own copy
// until the process ends.  In host code this name stands for a host-side
// symbol, not a device address: kernels receive ibpm_ctx::scratch, which
// ibpm_ctx_create takes from hipGetSymbolAddress with the device current.
static __device__ __align__(256) double scratch[IBPM_NOALLOC_MAX_POINTS];

struct ibpm_group;

namespace {

struct DevTables {
    DstPlan plan{};
    double2* tw = nullptr;
    double* sn = nullptr;
    double2* wm = nullptr;
};

struct CoarsifyRange {
    int i0, ic0, ci, j0, jc0, cj;
};

// CD/CL history of queued steps: a device buffer written by k_forces and a
// pinned host mirror it is copied to when the steps are waited for.  It
// grows geometrically between waits.  The buffers it outgrows are released
// only with their owner: hipFree and hipHostFree wait for the whole device,
// which would serialise problems that step on their own streams.
struct HistBuf {
    double* d = nullptr;
    double* h = nullptr;
    size_t n = 0;  // doubles in d and in h
    std::vector<void*> old_d, old_h;
    hipError_t reserve(size_t want)
    {
        if (want <= n) return hipSuccess;
        const size_t cap = std::max({want, 2 * n, (size_t)256});
        double *nd = nullptr, *nh = nullptr;
        hipError_t e = hipMalloc((void**)&nd, sizeof(double) * cap);
        if (e != hipSuccess) return e;
        e = hipHostMalloc((void**)&nh, sizeof(double) * cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            old_d.push_back(nd);
            return e;
        }
        if (d) old_d.push_back(d);
        if (h) old_h.push_back(h);
        d = nd;
        h = nh;
        n = cap;
        return hipSuccess;
    }
    void release()  // the owner's stream is drained
    {
        if (d) old_d.push_back(d);
        if (h) old_h.push_back(h);
        for (void* p : old_d) (void)hipFree(p);
        for (void* p : old_h) (void)hipHostFree(p);
        old_d.clear();
        old_h.clear();
        d = h = nullptr;
        n = 0;
    }
};

// E' of the current body points as a segmented gather: for every touched flux
// cell (ordered), the contributions (F̃b index, weight) in body-point order
struct RegLists {
    std::vector<int> cell, off, ef;
    std::vector<double> ew;
};

}  // namespace

struct ibpm_prob {
    ibpm_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    GridDev g{};
    int mg = 1;
    double Re = 0, dt = 0, beta[2] = {1.5, -0.5};
    double offx = 0, offy = 0;
    int nbodies = 0, nbt = 0, nf = 0, motion = 0;
    std::vector<int> nb, boff;
    std::vector<double> xb, ds;
    std::vector<int> body_idx;  // nbt x 2, Julia 1-based values
    std::vector<double> weight; // [(2k + c) * 169 + l + 13 m]
    std::vector<double> binv;   // nf x nf column-major
    double b_asym = 0.0, setup_s = 0.0;
    // q0 holds a per-level uniform value per component (inertial base flux,
    // timestepping.jl:262-269): kernels then take the values, not the array
    bool q0_uniform = true;
    double q0vx = 0.0, q0vy = 0.0;  // q0 = (q0vx, q0vy) hc on every level
    // LinearizedIBProblem (problem-types.jl:47-87): base flux QB (avg_flux of
    // the base state) and base circulation ΓB on the device; q0 stays zero
    bool lin = false;
    double *QB = nullptr, *GB = nullptr;
    int ytile = 4;         // y-pass columns per workgroup, all-level launches (and batches of fields)
    int ytile_single = 4;  // same, one-level launches (vort2flux chain)
    int pipe = 1;          // IBPM_OPT_PIPE
    int ysplit = 1;        // IBPM_OPT_YSPLIT
    int xreuse = 0;        // IBPM_OPT_XREUSE (default by grid size, set at creation)
    int xwaves = 4;        // lines per work-group of the plain x-passes over all levels (or a batch)

    // device state (IBState) and work arrays
    double *q = nullptr, *q0 = nullptr, *G = nullptr, *Gs = nullptr, *psi = nullptr;
    double* nl[2] = {nullptr, nullptr};
    int nl1 = 0, nl2 = 0;  // buffers holding state.nonlin[1], state.nonlin[2]
    double *qs1 = nullptr, *T1 = nullptr, *T2 = nullptr, *Gtmp = nullptr, *qreg = nullptr;
    // columns nx/4 and 3nx/4 of the last ψ level of the vort2flux! chain, as
    // (L, R) pairs per row: the get_bc! input of the next finer level
    double* edge = nullptr;
    double *Fb = nullptr, *Ftmp = nullptr, *fbv = nullptr, *CD = nullptr, *CL = nullptr, *scal = nullptr;
    double *BinvT = nullptr, *w = nullptr, *ds_d = nullptr;
    int *bidx = nullptr, *boff_d = nullptr;
    int *sc_cell = nullptr, *sc_off = nullptr, *sc_f = nullptr;
    double* sc_w = nullptr;
    int sc_ncell = 0;
    int rb_i0 = 2, rb_i1 = 1, rb_j0 = 2, rb_j1 = 1;  // rot! bounding box of E' support
    double *cx = nullptr, *cy = nullptr;
    DevTables tx, ty;
    std::vector<CoarsifyRange> crange;
    RegLists reg;
    // MotionFunction (lab-frame moving bodies): reference points x0 and point
    // velocities ub (nbt x 2 column-major, like xb); device output of B_times!
    // and pinned staging of its input and output for the host-side CG; the
    // last CG solve's products, final |r| and tolerance
    std::vector<double> x0, ub;
    double* Fw = nullptr;
    double* pin = nullptr;
    // B assembled every step in batches of bk unit-force fields: E' e_f (bq,
    // also the level-1 flux out), rot! (bgt), transform scratch (bt1), Γ and ψ
    // of every level (bG, bpsi: level l of field b at (l bk + b) nG), B on the
    // device (nf x nf column-major) and on the host
    int bk = 0;
    double *bq = nullptr, *bgt = nullptr, *bt1 = nullptr, *bG = nullptr, *bpsi = nullptr, *Bd = nullptr;
    double* bedge = nullptr;  // edge columns of every field of a batch
    std::vector<double> Bh;
    int cg_iters = 0;
    double cg_res = 0.0, cg_tol = 0.0;
    HistBuf hb;               // CD/CL history storage (device and pinned host mirror)
    double* hist = nullptr;   // = hb.d: CD history [hist_steps x nbodies], then CL history
    size_t hist_steps = 0;
    int pending = 0;          // steps enqueued by ibpm_advance_async, not yet waited for
    int* slot = nullptr;      // device: history entry of the next step (advanced by k_forces)
    bool async_err = false;   // a graph launch failed (reported by the next wait)
    ibpm_group* group = nullptr;  // set while the problem's state lives in a group's arrays
    // q_stale: P->q does not hold the state's q; curl! of P->psi (all levels)
    // gives it.  The FP64 step leaves q so, since the only reader of q inside a
    // step, the next step's nonlinear!, forms the fluxes from psi itself;
    // every other reader forms q first (ensure_q).
    bool q_stale = false;
    // 0-based lines of the level-1 ψ that E of the trial flux reads (curl! of
    // its window fluxes), or -1 when a window reaches a flux on the boundary
    int tlo = -1, thi = -1;
    // HIP graphs of one step (static bodies): replayed instead of ~54 launches.
    // A step's launches depend on the buffer roles only (Γ/Γs and the nonlin
    // pair alternate), so a graph is kept per role; anything else a captured
    // launch holds by value invalidates them (graphs_clear).
    bool graphs = true;
    struct StepGraph {
        double* G;
        int out, nl2;
        bool qs;  // q_stale at the start of the step: its nonlinear! reads psi
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> gcache;
    std::vector<void*> allocs;
    // Mixed precision (IBPM_OPT_PRECISION): -1 the FP64 transforms; MX_FP32 or
    // MX_BF16X3 the MFMA sine-matrix transforms.  Padded sizes P1, P2 (n1, n2
    // rounded up to 64), the sine matrices of both dimensions (FP32 and bf16
    // high/low parts, P x P), two padded FP32 fields per level, the per-level
    // operand scales
    int prec = -1;
    int mx_sweeps = 2;  // refinement sweeps of every mixed-precision solve (IBPM_OPT_MX_SWEEPS)
    int P1 = 0, P2 = 0;
    float *mx_sf1 = nullptr, *mx_sf2 = nullptr, *mx_buf[2] = {nullptr, nullptr};
    float* mx_part = nullptr;  // split-K partial products (MX_SPLIT_MAX x mg fields)
    __bf16 *mx_sh1 = nullptr, *mx_sl1 = nullptr, *mx_sh2 = nullptr, *mx_sl2 = nullptr;
    double* mx_lscale = nullptr;
    // tangent-linear / adjoint work space (allocated by their first call)
    struct AdjWork* adj = nullptr;
    // in-step timing (IBPM_OPT_PROFILE): 1 the all-level y-pass launches, 2
    // the one-level y-passes with boundary terms of the vort2flux! chain;
    // event pairs recorded on the problem's stream, summed after each call
    int prof = 0;
    std::vector<hipEvent_t> pev;
    size_t pev_used = 0;
    double prof_ms = 0.0;
    long prof_n = 0;
};

static void graphs_clear(ibpm_prob* P);

// Device buffers of the tangent-linear and adjoint steps (MovingGrid bodies).
// Per level: nG for circulation-like fields, nq for fluxes, nbc for boundary
// vectors.  snap: the binomial checkpoints, (Γ, nonlin[2], ψ) each.
struct AdjWork {
    int nsteps_cap = 0, nsnap = 0;
    double *dG = nullptr, *dGs = nullptr, *dN[2] = {nullptr, nullptr}, *tG = nullptr, *dpsi = nullptr;
    double *dq = nullptr, *dq0 = nullptr, *dqs1 = nullptr, *Q = nullptr, *dQ = nullptr, *fq = nullptr;
    double *F1 = nullptr, *F2 = nullptr, *BinvC = nullptr, *bc = nullptr, *bcb = nullptr, *tG1 = nullptr;
    double *Gb = nullptr, *Gbo = nullptr, *Nb1 = nullptr, *Nb0 = nullptr, *Rb = nullptr, *psib = nullptr;
    double *qb = nullptr, *q0b = nullptr, *qs1b = nullptr, *part = nullptr;
    double* sums = nullptr;  // per step: TLM Σ δF̃b_x, Σ δF̃b_y; adjoint Ūx0, Ūy0, Ω̄
    // the state's q0, F̃b, CD and CL at the call, put back at its end
    double* aux = nullptr;
    bool q0_uniform = true;
    double q0vx = 0.0, q0vy = 0.0;
    std::vector<double*> snap;
    std::vector<void*> allocs;
};

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static int fail(ibpm_ctx* c, int code, const std::string& m)
{
    if (c) c->err = m;
    return code;
}

#define HIPTRY(ctx, expr)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail((ctx), IBPM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

// Entry points that replace buffers or captured step graphs refuse to run
// while steps enqueued by ibpm_advance_async are still in flight.
static int check_idle(ibpm_prob* P)
{
    if (P->pending > 0)
        return fail(P->ctx, IBPM_E_ARG, "steps enqueued by ibpm_advance_async are not waited for");
    return IBPM_OK;
}

template <typename T>
static hipError_t dalloc(ibpm_prob* P, T** p, size_t n)
{
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (n > 0 ? n : 1));
    if (e == hipSuccess) {
        P->allocs.push_back(v);
        *p = (T*)v;
    }
    return e;
}

// ---------------------------------------------------------------------------
// host-side setup pieces
// ---------------------------------------------------------------------------
static const long double PI_L = 3.14159265358979323846264338327950288L;

// DST-I plan for line length M-1.  Returns false when M is unsupported.
static bool make_plan(int M, DstPlan& pl, std::vector<double2>& tw, std::vector<double>& sn, std::vector<double2>& wm)
{
    pl = DstPlan{};
    pl.M = M;
    if (M < 2) return false;
    bool fft = (M % 2 == 0) && (M <= MAXM);
    int P = M / 2;
    std::vector<int> rad;
    if (fft) {
        int m = P, e2 = 0;
        while (m % 2 == 0) { m /= 2; e2++; }
        int e3 = 0, e5 = 0;
        while (m % 3 == 0) { m /= 3; e3++; }
        while (m % 5 == 0) { m /= 5; e5++; }
        if (m != 1) fft = false;
        else {
            while (e2 >= 3) { rad.push_back(8); e2 -= 3; }
            if (e2 == 2) rad.push_back(4);
            if (e2 == 1) rad.push_back(2);
            for (int k = 0; k < e5; k++) rad.push_back(5);
            for (int k = 0; k < e3; k++) rad.push_back(3);
        }
    }
    if (fft) {
        pl.kind = 0;
        pl.P = P;
        pl.nst = (int)rad.size();
        if (pl.nst > 8) return false;
        pl.rad = 0;
        for (int k = 0; k < pl.nst; k++) pl.rad |= (unsigned)rad[k] << (4 * k);
        tw.resize(P);
        for (int e = 0; e < P; e++) {
            long double a = -2.0L * PI_L * (long double)e / (long double)P;
            tw[e] = make_double2((double)cosl(a), (double)sinl(a));
        }
        sn.resize(P + 1);
        for (int j = 0; j <= P; j++) sn[j] = (double)sinl(PI_L * (long double)j / (long double)M);
        wm.resize(P);
        for (int k = 0; k < P; k++) {
            long double a = -2.0L * PI_L * (long double)k / (long double)M;
            wm[k] = make_double2((double)cosl(a), (double)sinl(a));
        }
        return true;
    }
    if (M - 1 > MAXM) return false;
    pl.kind = 1;
    pl.P = 0;
    pl.nst = 0;
    tw.assign(1, make_double2(0, 0));
    wm.assign(1, make_double2(0, 0));
    sn.resize(2 * M);
    for (int t = 0; t < 2 * M; t++) sn[t] = (double)sinl(PI_L * (long double)t / (long double)M);
    return true;
}

// δh (interface-coupling.jl:6-62): Yang-3 smoothed delta times h.
static double delta_h(double rf, double rb, double dr)
{
    double r = std::fabs(rf - rb);
    double r1 = r / dr, r2 = r1 * r1, r3 = r2 * r1, r4 = r3 * r1;
    double del_h;
    if (r1 <= 1.0) {
        double a5 = std::asin((1.0 / 2.0) * std::sqrt(3.0) * (2.0 * r1 - 1.0));
        double a8 = std::sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
        del_h = 4.166666667e-2 * r4 + (-0.1388888889 + 3.472222222e-2 * a8) * r3 +
                (-7.121664902e-2 - 5.208333333e-2 * a8 + 0.2405626122 * a5) * r2 +
                (-.2405626122 * a5 - 0.3792313933 + 0.1012731481 * a8) * r1 + 8.0187537413e-2 * a5 -
                4.195601852e-2 * a8 + 0.6485698427;
    } else if (r1 <= 2.0) {
        double a6 = std::asin((1.0 / 2.0) * std::sqrt(3.0) * (-3.0 + 2.0 * r1));
        double a9 = std::sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
        del_h = -6.250000000e-2 * r4 + (0.4861111111 - 1.736111111e-2 * a9) * r3 +
                (-1.143175026 + 7.812500000e-2 * a9 - 0.1202813061 * a6) * r2 +
                (0.8751991178 + 0.3608439183 * a6 - 0.1548032407 * a9) * r1 - 0.2806563809 * a6 +
                8.22848104e-3 + 0.1150173611 * a9;
    } else if (r1 <= 3.0) {
        double a1 = std::asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * std::sqrt(3.0));
        double a7 = std::sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
        del_h = 2.083333333e-2 * r4 + (3.472222222e-3 * a7 - 0.2638888889) * r3 +
                (1.214391675 - 2.604166667e-2 * a7 + 2.405626122e-2 * a1) * r2 +
                (-0.1202813061 * a1 - 2.449273192 + 7.262731481e-2 * a7) * r1 + 0.1523563211 * a1 +
                1.843201677 - 7.306134259e-2 * a7;
    } else {
        del_h = 0.0;
    }
    return del_h;
}

// ---------------------------------------------------------------------------
// operator schedules on the device
// ---------------------------------------------------------------------------
static XArgs xargs(ibpm_prob* P)
{
    XArgs a{};
    a.lstride = P->g.nG;
    a.mg = P->mg;
    a.pl = P->tx.plan;
    a.g = P->g;
    a.zlo = 1;
    a.zhi = P->g.n2;
    a.waves = P->xwaves;
    return a;
}

static void set_nl_q0(const ibpm_prob* P, NLArgs& na)
{
    na.q0_uniform = P->q0_uniform ? 1 : 0;
    na.q0vx = P->q0vx;
    na.q0vy = P->q0vy;
    na.QB = P->lin ? P->QB : nullptr;
    na.GB = P->lin ? P->GB : nullptr;
}

// IBPM_OPT_XREUSE by default: on for grids of 2^19 cells and more, where the
// x-transform the step's vort2flux! chains skip is worth more than the second
// source their y-passes then read (measured: +4 % at 1024 x 1024, -3 to -5 %
// at 800 x 400 and below)
static int xreuse_default(int nx, int ny) { return (long)nx * ny >= (1L << 19) ? 1 : 0; }

static YArgs yargs(ibpm_prob* P, int op, int lev0, int nlev, double* buf)
{
    YArgs a{};
    a.buf = buf;
    a.lstride = P->g.nG;
    a.nlev = nlev;
    a.lev0 = lev0;
    a.op = op;
    a.cx = P->cx;
    a.cy = P->cy;
    a.dt = P->dt;
    a.Re = P->Re;
    a.scale = 1.0 / (double)(4L * P->g.nx * P->g.ny);
    a.sinx = P->tx.sn;  // sin(pi m / nx), m <= nx/2, leads either plan kind's table
    a.bcfac = 1.0;
    a.pl = P->ty.plan;
    a.g = P->g;
    a.pipe = P->pipe;
    a.split = P->ysplit;
    return a;
}

// setup_reg (interface-coupling.jl:72-96) for the points in P->xb: body_idx,
// the 13 x 13 weights, the E' segmented-gather lists (flux cell, F̃b index,
// weight ordered by cell, then by point) and the rot! box of the E' support.
// Returns false when a point is closer than 6 cells to the domain edge.
static bool setup_reg_host(ibpm_prob* P)
{
    const GridDev& g = P->g;
    const int nx = g.nx, ny = g.ny, nbt = P->nbt;
    const double h = g.h;
    P->body_idx.resize(2 * (size_t)nbt);
    for (int k = 0; k < nbt; k++) {
        P->body_idx[k] = (int)std::floor((P->xb[k] + P->offx) / h);
        P->body_idx[nbt + k] = (int)std::floor((P->xb[nbt + k] + P->offy) / h);
        int bx = P->body_idx[k], by = P->body_idx[nbt + k];
        if (bx - 6 < 1 || bx + 6 > nx || by - 6 < 1 || by + 6 > ny) return false;
    }
    P->weight.assign((size_t)2 * nbt * 169, 0.0);
    for (int k = 0; k < nbt; k++) {
        double xk = P->xb[k], yk = P->xb[nbt + k];
        int bx = P->body_idx[k], by = P->body_idx[nbt + k];
        for (int m = 0; m < 13; m++) {
            double y = h * (double)(by - 1 + (m - 6)) - P->offy;
            for (int l = 0; l < 13; l++) {
                double x = h * (double)(bx - 1 + (l - 6)) - P->offx;
                P->weight[(size_t)(2 * k) * 169 + l + 13 * m] = delta_h(x, xk, h) * delta_h(y + h / 2, yk, h);
                P->weight[(size_t)(2 * k + 1) * 169 + l + 13 * m] = delta_h(x + h / 2, xk, h) * delta_h(y, yk, h);
            }
        }
    }
    struct Ent { long p; int f; double w; };
    std::vector<Ent> ents;
    ents.reserve((size_t)2 * nbt * 169);
    int ib0 = nx + 1, ib1 = 0, jb0 = ny + 2, jb1 = 0;
    for (int c = 0; c < 2; c++)
        for (int k = 0; k < nbt; k++) {
            int bx = P->body_idx[k], by = P->body_idx[nbt + k];
            for (int m = 0; m < 13; m++)
                for (int l = 0; l < 13; l++) {
                    int i = bx + l - 6, j = by + m - 6;
                    long p = (c == 0) ? (i - 1) + (long)(nx + 1) * (j - 1) : g.nu + (i - 1) + (long)nx * (j - 1);
                    ents.push_back({p, k + c * nbt, P->weight[(size_t)(2 * k + c) * 169 + l + 13 * m]});
                    // rot! outputs touched: x-flux (i,j) -> Γ(i,j),(i,j+1); y-flux -> Γ(i,j),(i+1,j)
                    ib0 = std::min(ib0, i); jb0 = std::min(jb0, j);
                    ib1 = std::max(ib1, i + 1); jb1 = std::max(jb1, j + 1);
                }
        }
    std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) { return a.p < b.p; });
    RegLists& R = P->reg;
    R.cell.clear(); R.off.clear(); R.ef.clear(); R.ew.clear();
    for (size_t e = 0; e < ents.size(); e++) {
        if (e == 0 || ents[e].p != ents[e - 1].p) {
            R.cell.push_back((int)ents[e].p);
            R.off.push_back((int)e);
        }
        R.ef.push_back(ents[e].f);
        R.ew.push_back(ents[e].w);
    }
    R.off.push_back((int)ents.size());
    P->sc_ncell = (int)R.cell.size();
    // the windows' fluxes (i, j), i = ib0..ib1-1, j = jb0..jb1-1, read ψ rows
    // j-1 and j (1-based); all interior: curl! of them needs no boundary value
    if (ib0 >= 2 && ib1 - 1 <= nx - 1 && jb0 >= 2 && jb1 - 1 <= ny - 1) {
        P->tlo = jb0 - 2;
        P->thi = jb1 - 2;
    } else {
        P->tlo = P->thi = -1;
    }
    P->rb_i0 = std::max(ib0, 2); P->rb_i1 = std::min(ib1, nx);
    P->rb_j0 = std::max(jb0, 2); P->rb_j1 = std::min(jb1, ny);
    return true;
}

// E tables of setup_reg_host to the device, on the problem's stream.  The
// device arrays hold the largest lists (every window entry its own cell).
static hipError_t upload_reg(ibpm_prob* P)
{
    hipStream_t s = P->stream;
    const RegLists& R = P->reg;
    auto cp = [&](void* d, const void* h, size_t bytes) {
        return bytes ? hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    hipError_t e = cp(P->w, P->weight.data(), sizeof(double) * P->weight.size());
    if (e == hipSuccess) e = cp(P->bidx, P->body_idx.data(), sizeof(int) * P->body_idx.size());
    if (e == hipSuccess) e = cp(P->sc_cell, R.cell.data(), sizeof(int) * R.cell.size());
    if (e == hipSuccess) e = cp(P->sc_off, R.off.data(), sizeof(int) * R.off.size());
    if (e == hipSuccess) e = cp(P->sc_f, R.ef.data(), sizeof(int) * R.ef.size());
    if (e == hipSuccess) e = cp(P->sc_w, R.ew.data(), sizeof(double) * R.ew.size());
    return e;
}

// x = L b for one level (lev 0-based), L one of Δinv / A / Ainv (laplacian.jl:24-76),
// with an optional subtract epilogue.  tmp: the level's transform scratch
// (nG doubles, every one written by the first pass), by default T1's level.
// xmode X_ROT: b is a level-1 flux field and the operand is rot!(b), formed in
// the first x-pass
static void lap_level(ibpm_prob* P, int op, int lev, const double* b, double* x, bool sub, int zlo = 1,
                      int zhi = 1 << 30, double* tmp = nullptr, int xmode = X_PLAIN)
{
    hipStream_t s = P->stream;
    if (!tmp) tmp = P->T1 + (long)lev * P->g.nG;
    XArgs a = xargs(P);
    a.src = b;
    a.dst = tmp;
    a.nlev = 1;
    a.lev0 = lev;
    YArgs y = yargs(P, op, lev, 1, tmp);
    zhi = std::min(zhi, P->g.n2);
    if (zlo > 1 || zhi < P->g.n2) {
        // input lines outside zlo..zhi are zero: only those lines are
        // transformed, and the y-pass reads no others
        a.lsel = y.rsel = 1;
        a.llo = y.rlo = zlo - 1;
        a.lhi = y.rhi = zhi - 1;
    }
    launch_xpass(xmode, a, s);
    launch_ypass(y, P->ytile_single, s);
    XArgs c = xargs(P);
    c.src = tmp;
    c.dst = x;
    c.nlev = 1;
    c.lev0 = lev;
    launch_xpass(sub ? X_SUB : X_PLAIN, c, s);
}

static void timed_ypass(ibpm_prob* P, const YArgs& y, int tile, int kind);

// vort2flux! (curl.jl:133-165).  Gb is mutated on coarse levels.  When
// `level1_only`, only q on level 1 is produced (into qout), which is all the
// trial stage and B_times! read of it.
// what vort2flux_dev stores of q = curl!(ψ): every level, level 1 only, or
// nothing (the caller forms the fluxes it needs from ψ).  CURL_TRIAL: nothing,
// and of ψ only what the trial's E and the chain itself read: the level-1
// lines tlo..thi, and of each coarser level the lines ny/4 .. 3ny/4 (1-based)
// the next finer level's boundary values come from, whose edge columns are
// kept and whose rows ny/4 and 3ny/4 are stored (every line when tlo < 0).
enum { CURL_ALL = 0, CURL_LEVEL1 = 1, CURL_NONE = 2, CURL_TRIAL = 3 };

// yhat (optional): a field whose x-transform is Gb on every line coarsify!
// leaves alone: Gb = X(yhat) there, X the x-DST-I (Ainv's last pass wrote Gb
// from yhat), so X(Gb) = X(X(yhat)) = 2 nx yhat on those lines.  Only the
// lines coarsify! rewrites (the centre rows of levels 2..L) are then
// x-transformed; the y-pass reads the others as 2 nx yhat.  ψ then differs
// from the full transform's by rounding.  ysub0 (optional, with yhat): level
// 1 of Gb is X(yhat) - X(ysub0) instead, read as 2 nx (yhat - ysub0).
static void vort2flux_dev(ibpm_prob* P, double* Gb, double* psib, double* qout, int curl,
                          const double* yhat = nullptr, const double* ysub0 = nullptr)
{
    hipStream_t s = P->stream;
    const long nG = P->g.nG;
    const int L = P->mg;
    for (int lev = 2; lev <= L; lev++) {
        CoarsifyArgs c{};
        const CoarsifyRange& r = P->crange[0];
        c.Gc = Gb + (long)(lev - 1) * nG;
        c.G = Gb + (long)(lev - 2) * nG;
        c.i0 = r.i0; c.ic0 = r.ic0; c.ci = r.ci; c.j0 = r.j0; c.jc0 = r.jc0; c.cj = r.cj;
        c.g = P->g;
        launch_coarsify(c, s);
    }
    // Δinv of every level, coarse to fine.  The x-transform of Γ does not
    // depend on the chain, so all levels take it in one launch; the boundary
    // terms of apply_bc! (from ψ of the level above) enter, x-transformed, in
    // the y-pass; the inverse x-pass of a level stores the columns the next
    // finer level's get_bc! reads.
    XArgs a = xargs(P);
    // 0-based lines of the centre rows coarsify! writes on levels 2..L
    const int clo = P->crange[0].jc0 - 1, chi = P->crange[0].jc0 + P->crange[0].cj - 2;
    if (!yhat) {
        a.src = Gb; a.dst = P->T1; a.nlev = L; a.lev0 = 0;
        launch_xpass(X_PLAIN, a, s);
    } else if (L > 1 && chi >= clo) {
        a.src = Gb + nG; a.dst = P->T1 + nG; a.nlev = L - 1; a.lev0 = 1;
        a.lsel = 1; a.llo = clo; a.lhi = chi;
        launch_xpass(X_PLAIN, a, s);
    }
    const bool trial = curl == CURL_TRIAL && P->tlo >= 0;
    for (int l = L - 1; l >= 0; l--) {
        YArgs y = yargs(P, Y_DINV, l, 1, P->T1 + (long)l * nG);
        if (yhat) {
            y.ysrc = yhat + (long)l * nG;
            y.afac = 2.0 * (double)P->g.nx;
            y.clo = l > 0 ? clo : 1;  // level 1: every line from yhat
            y.chi = l > 0 ? chi : 0;
            if (l == 0) y.ysub = ysub0;
        }
        if (l < L - 1) {
            y.bcr = psib + (long)(l + 1) * nG;
            y.edge = P->edge;
        }
        if (trial) {  // the rows the trial's x-pass of this level reads (below)
            y.wsel = 1;
            y.wlo = l > 0 ? P->g.ny / 4 - 1 : P->tlo;
            y.whi = l > 0 ? 3 * P->g.ny / 4 - 1 : P->thi;
        }
        // in-step timing (kind 2) of the launches that read and write the whole level
        timed_ypass(P, y, P->ytile_single, (l < L - 1 && !trial) ? 2 : -1);
        XArgs c = xargs(P);
        c.src = P->T1 + (long)l * nG; c.dst = psib + (long)l * nG; c.nlev = 1; c.lev0 = l;
        if (l > 0) c.edge = P->edge;
        if (trial) {
            c.lsel = 1;
            if (l > 0) {
                c.llo = c.keep0 = P->g.ny / 4 - 1;
                c.lhi = c.keep1 = 3 * P->g.ny / 4 - 1;
                c.keep = 1;
            } else {
                c.llo = P->tlo;
                c.lhi = P->thi;
            }
        }
        launch_xpass(X_PLAIN, c, s);
    }
    if (qout == P->q) P->q_stale = curl != CURL_ALL;
    if (curl == CURL_NONE || curl == CURL_TRIAL) return;
    CurlArgs ca{};
    ca.psi = psib;
    ca.q = qout;
    ca.qstride = P->g.nq;
    ca.nlev = curl == CURL_LEVEL1 ? 1 : L;
    ca.lev0 = 0;
    ca.mg = L;
    ca.g = P->g;
    launch_curl(ca, s);
}

// the state's q formed from its ψ when the last step left it unformed
static void ensure_q(ibpm_prob* P)
{
    if (!P->q_stale) return;
    CurlArgs ca{};
    ca.psi = P->psi;
    ca.q = P->q;
    ca.qstride = P->g.nq;
    ca.nlev = P->mg;
    ca.lev0 = 0;
    ca.mg = P->mg;
    ca.g = P->g;
    launch_curl(ca, P->stream);
    P->q_stale = false;
}

// where a step's nonlinear! reads the state's fluxes: q, or ψ when the last
// step left q unformed (the linearized product reads q, which is formed first)
static void nl_flux_source(ibpm_prob* P, NLArgs& na)
{
    if (!P->q_stale) return;
    if (P->lin) ensure_q(P);
    else na.psi = P->psi;
}

// Γout[:,1] = Ainv[1] * rot(reg(f)) -- the body-force projection of
// project_circ! (timestepping.jl:217-218) and of B_times! (laplacian.jl:162).
// sub: Γout[:,1] -= result instead of storing it.
// fa: update_stress! of F̃b = f, run in the scatter's launch (null: none)
static void project_forces(ibpm_prob* P, const double* f, double* Gout, bool sub, const ForceArgs* fa = nullptr)
{
    hipStream_t s = P->stream;
    ScatterArgs sa{};
    sa.cell = P->sc_cell; sa.off = P->sc_off; sa.ent_f = P->sc_f; sa.ent_w = P->sc_w;
    sa.fb = f; sa.q = P->qreg; sa.ncell = P->sc_ncell;
    if (fa) launch_scatter_forces(sa, *fa, s);
    else launch_scatter(sa, s);
    // rot!(qreg) is formed in Ainv's first x-pass; it is zero outside the rows
    // rb_j0 - 1 .. rb_j1 - 1 of the rot! box (qreg is zero outside the E' support)
    lap_level(P, Y_AINV, 0, P->qreg, Gout, sub, P->rb_j0 - 1, P->rb_j1 - 1, nullptr, X_ROT);
}

// A y-pass of the step with `tile` columns per work-group, bracketed by a pair
// of events on the problem's stream when in-step profiling of `kind` is on.
static void timed_ypass(ibpm_prob* P, const YArgs& y, int tile, int kind)
{
    hipEvent_t e[2] = {nullptr, nullptr};
    if (P->prof == kind) {
        for (int k = 0; k < 2; k++) {
            if (P->pev_used == P->pev.size()) {
                hipEvent_t ev;  // timing only: no system-scope fence (no cache writeback per record)
                if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) break;
                P->pev.push_back(ev);
            }
            e[k] = P->pev[P->pev_used++];
        }
    }
    if (e[1]) (void)hipEventRecord(e[0], P->stream);
    launch_ypass(y, tile, P->stream);
    if (e[1]) (void)hipEventRecord(e[1], P->stream);
}

// fold the recorded event pairs into the running sums (stream already drained)
static void collect_profile(ibpm_prob* P)
{
    for (size_t k = 0; k + 1 < P->pev_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, P->pev[k], P->pev[k + 1]) == hipSuccess) {
            P->prof_ms += ms;
            P->prof_n++;
        }
    }
    P->pev_used = 0;
}

static void trial_dev(ibpm_prob* P);
static void trial_flux_from_psi(ibpm_prob* P, GatherArgs& ga);

// ---------------------------------------------------------------------------
// Mixed precision: the Laplacian inverses of the step as dense sine-matrix
// transforms on the MFMA units (mixed.hip), in increment form.  Ainv and
// Δinv act on residuals computed in FP64 from the current state, and their
// results are added to FP64 fields:
//   trial:     Γs = Γ + Ainv(A Γ - Ainv^{-1} Γ + β1 dt N1 + β2 dt N0 + rhsbc),
//              with A Γ - Ainv^{-1} Γ = -2 qc L Γ by the 5-point stencil;
//   vort2flux: per level, coarse to fine, ψ += Δinv(Γ + apply_bc!(get_bc!(ψ of
//              the coarser level)) - L ψ), ψ starting from the state's ψ;
//   projection: Γ[:,1] -= Ainv rot(E' F̃b).
// In exact arithmetic these are the reference's operators (A and Ainv are
// diagonal in the same sine basis; vort2flux! is linear and the state's ψ
// solves it for the state's Γ), so the low-precision transforms only see the
// per-step change of each field, and their rounding is relative to it.  A
// second sweep of each solve (the default) applies the transforms once more
// to the FP64 residual the first one left, squaring the relative error.
// ---------------------------------------------------------------------------
static uint16_t bf16_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// P x P sine matrix of RODFT00 for lines of n points, zero outside n x n, as
// FP32 and as bf16 high and low parts
static void sine_matrix(int n, int Pn, std::vector<float>& f, std::vector<uint16_t>& hi, std::vector<uint16_t>& lo)
{
    const int N2 = 2 * (n + 1);
    std::vector<double> tab(N2);
    for (int t = 0; t < N2; t++) tab[t] = (double)(2.0L * sinl(PI_L * (long double)t / (long double)(n + 1)));
    f.assign((size_t)Pn * Pn, 0.f);
    hi.assign((size_t)Pn * Pn, 0);
    lo.assign((size_t)Pn * Pn, 0);
    for (int k = 0; k < n; k++)
        for (int i = 0; i < n; i++) {
            const double v = tab[((long)(k + 1) * (i + 1)) % N2];
            const size_t e = (size_t)k * Pn + i;
            f[e] = (float)v;
            const uint16_t h = bf16_bits((float)v);
            hi[e] = h;
            lo[e] = bf16_bits((float)(v - (double)bf16_value(h)));
        }
}

// device tables and buffers of the mixed form (once per problem)
static int mx_setup(ibpm_prob* P)
{
    if (P->mx_sf1) return IBPM_OK;
    const GridDev& g = P->g;
    P->P1 = (g.n1 + 63) / 64 * 64;
    P->P2 = (g.n2 + 63) / 64 * 64;
    const size_t a1 = (size_t)P->P1 * P->P1, a2 = (size_t)P->P2 * P->P2, fl = (size_t)P->P1 * P->P2 * P->mg;
    HIPTRY(P->ctx, dalloc(P, &P->mx_sf1, a1));
    HIPTRY(P->ctx, dalloc(P, &P->mx_sf2, a2));
    HIPTRY(P->ctx, dalloc(P, (uint16_t**)&P->mx_sh1, a1));
    HIPTRY(P->ctx, dalloc(P, (uint16_t**)&P->mx_sl1, a1));
    HIPTRY(P->ctx, dalloc(P, (uint16_t**)&P->mx_sh2, a2));
    HIPTRY(P->ctx, dalloc(P, (uint16_t**)&P->mx_sl2, a2));
    HIPTRY(P->ctx, dalloc(P, &P->mx_buf[0], fl));
    HIPTRY(P->ctx, dalloc(P, &P->mx_buf[1], fl));
    HIPTRY(P->ctx, dalloc(P, &P->mx_part, (size_t)MX_SPLIT_MAX * fl));
    HIPTRY(P->ctx, dalloc(P, &P->mx_lscale, (size_t)P->mg));
    std::vector<float> f;
    std::vector<uint16_t> hi, lo;
    sine_matrix(g.n1, P->P1, f, hi, lo);
    HIPTRY(P->ctx, hipMemcpy(P->mx_sf1, f.data(), sizeof(float) * a1, hipMemcpyHostToDevice));
    HIPTRY(P->ctx, hipMemcpy(P->mx_sh1, hi.data(), 2 * a1, hipMemcpyHostToDevice));
    HIPTRY(P->ctx, hipMemcpy(P->mx_sl1, lo.data(), 2 * a1, hipMemcpyHostToDevice));
    sine_matrix(g.n2, P->P2, f, hi, lo);
    HIPTRY(P->ctx, hipMemcpy(P->mx_sf2, f.data(), sizeof(float) * a2, hipMemcpyHostToDevice));
    HIPTRY(P->ctx, hipMemcpy(P->mx_sh2, hi.data(), 2 * a2, hipMemcpyHostToDevice));
    HIPTRY(P->ctx, hipMemcpy(P->mx_sl2, lo.data(), 2 * a2, hipMemcpyHostToDevice));
    // per-level scale of the low-precision operands: Γ-like fields of level l
    // carry hc^2 = 4^l h^2, so 4^-l keeps every level's operands alike (exact)
    std::vector<double> ls(P->mg);
    for (int l = 0; l < P->mg; l++) ls[l] = std::ldexp(1.0, -2 * l);
    HIPTRY(P->ctx, hipMemcpy(P->mx_lscale, ls.data(), sizeof(double) * ls.size(), hipMemcpyHostToDevice));
    return IBPM_OK;
}

static MxResidArgs mx_rargs(ibpm_prob* P, int mode, int lev0)
{
    MxResidArgs r{};
    r.fstride = P->g.nG;
    r.out = P->mx_buf[0];
    r.ostride = (long)P->P1 * P->P2;
    r.P1 = P->P1;
    r.P2 = P->P2;
    r.mode = mode;
    r.lev0 = lev0;
    r.mg = P->mg;
    r.dt = P->dt;
    r.Re = P->Re;
    r.lscale = P->mx_lscale;
    r.g = P->g;
    return r;
}

// dst = base +/- L(residual) for nb levels from lev0, the residual already in
// mx_buf[0]: x-transform, y-transform with the FP64 eigenvalue scaling,
// y-transform, x-transform into the FP64 field
static void mx_apply(ibpm_prob* P, int op, int lev0, int nb, int epi, double* dst, const double* base)
{
    hipStream_t s = P->stream;
    MxGemmArgs a{};
    a.istride = a.ostride = (long)P->P1 * P->P2;
    a.fstride = P->g.nG;
    a.P1 = P->P1;
    a.P2 = P->P2;
    a.lev0 = lev0;
    a.op = op;
    a.cx = P->cx;
    a.cy = P->cy;
    a.dt = P->dt;
    a.Re = P->Re;
    a.scale = 1.0 / (double)(4L * P->g.nx * P->g.ny);
    a.lscale = P->mx_lscale;
    a.g = P->g;
    a.part = P->mx_part;
    a.partcap = (long)MX_SPLIT_MAX * P->P1 * P->P2 * P->mg;
    MxGemmArgs x = a, y = a;
    x.sf = P->mx_sf1; x.sh = P->mx_sh1; x.sl = P->mx_sl1;
    y.sf = P->mx_sf2; y.sh = P->mx_sh2; y.sl = P->mx_sl2;
    x.in = P->mx_buf[0]; x.out = P->mx_buf[1];
    launch_mx_gemm(P->prec, 0, MX_EPI_STORE, x, nb, s);
    y.in = P->mx_buf[1]; y.out = P->mx_buf[0];
    launch_mx_gemm(P->prec, 1, MX_EPI_SCALE, y, nb, s);
    y.in = P->mx_buf[0]; y.out = P->mx_buf[1];
    launch_mx_gemm(P->prec, 1, MX_EPI_STORE, y, nb, s);
    x.in = P->mx_buf[1]; x.dst = dst; x.base = base;
    launch_mx_gemm(P->prec, 0, epi, x, nb, s);
}

// vort2flux! in increment form on Γb (coarsified first): ψ updated in place,
// coarse to fine; then q of level 1 (level1_only) or of every level
static void mx_vort2flux(ibpm_prob* P, double* Gb, double* qout, bool level1_only)
{
    hipStream_t s = P->stream;
    const long nG = P->g.nG;
    const int L = P->mg;
    for (int lev = 2; lev <= L; lev++) {
        CoarsifyArgs c{};
        const CoarsifyRange& r = P->crange[0];
        c.Gc = Gb + (long)(lev - 1) * nG;
        c.G = Gb + (long)(lev - 2) * nG;
        c.i0 = r.i0; c.ic0 = r.ic0; c.ci = r.ci; c.j0 = r.j0; c.jc0 = r.jc0; c.cj = r.cj;
        c.g = P->g;
        launch_coarsify(c, s);
    }
    for (int l = L - 1; l >= 0; l--) {
        double* pl = P->psi + (long)l * nG;
        for (int sw = 0; sw < P->mx_sweeps; sw++) {
            MxResidArgs r = mx_rargs(P, MX_RES_V2F, l);
            r.G = Gb + (long)l * nG;
            r.psi = pl;
            launch_mx_resid(r, 1, s);
            mx_apply(P, MX_OP_DINV, l, 1, MX_EPI_ADD, pl, pl);
        }
    }
    CurlArgs ca{};
    ca.psi = P->psi;
    ca.q = qout;
    ca.qstride = P->g.nq;
    ca.nlev = level1_only ? 1 : L;
    ca.lev0 = 0;
    ca.mg = L;
    ca.g = P->g;
    launch_curl(ca, s);
}

// Γout[:,1] -/= Ainv[1] rot(reg(f)), mixed precision
static void mx_project_forces(ibpm_prob* P, const double* f, double* Gout, bool sub)
{
    hipStream_t s = P->stream;
    ScatterArgs sa{};
    sa.cell = P->sc_cell; sa.off = P->sc_off; sa.ent_f = P->sc_f; sa.ent_w = P->sc_w;
    sa.fb = f; sa.q = P->qreg; sa.ncell = P->sc_ncell;
    launch_scatter(sa, s);
    launch_rot(P->qreg, P->Gtmp, P->rb_i0, P->rb_i1, P->rb_j0, P->rb_j1, P->g, s);
    MxResidArgs r = mx_rargs(P, MX_RES_COPY, 0);
    r.G = P->Gtmp;
    launch_mx_resid(r, 1, s);
    if (P->mx_sweeps < 2) {
        mx_apply(P, MX_OP_AINV, 0, 1, sub ? MX_EPI_SUB : MX_EPI_ADD, Gout, sub ? Gout : nullptr);
        return;
    }
    // c = Ainv g in T1, refined by c += Ainv(g - Ainv^{-1} c), then applied
    double* c = P->T1;
    mx_apply(P, MX_OP_AINV, 0, 1, MX_EPI_ADD, c, nullptr);
    for (int sw = 1; sw < P->mx_sweeps; sw++) {
        MxResidArgs r2 = mx_rargs(P, MX_RES_PROJ2, 0);
        r2.G = P->Gtmp;
        r2.G2 = c;
        launch_mx_resid(r2, 1, s);
        mx_apply(P, MX_OP_AINV, 0, 1, MX_EPI_ADD, c, c);
    }
    if (sub) launch_mx_axpy(Gout, c, -1.0, P->g.nG, s);
    else launch_copy(Gout, c, P->g.nG, s);
}

// advance! with the mixed-precision transforms (same stages as step_body)
static void step_mixed(ibpm_prob* P)
{
    hipStream_t s = P->stream;
    const GridDev& g = P->g;
    const int L = P->mg;
    const int out = (P->nl1 != P->nl2) ? P->nl1 : 1 - P->nl2;
    NLArgs na{};
    na.q = P->q; na.q0 = P->q0; na.G = P->G; na.out = P->nl[out];
    nl_flux_source(P, na);
    na.nlev = L; na.lev0 = 0; na.mg = L; na.g = g;
    set_nl_q0(P, na);
    launch_nonlinear(na, s);
    MxResidArgs r = mx_rargs(P, MX_RES_TRIAL, 0);
    r.G = P->G;
    r.N1 = P->nl[out];
    r.N0 = P->nl[P->nl2];
    r.c1 = P->beta[0] * P->dt;
    r.c2 = P->beta[1] * P->dt;
    launch_mx_resid(r, L, s);
    mx_apply(P, MX_OP_AINV, 0, L, MX_EPI_ADD, P->Gs, P->G);
    for (int sw = 1; sw < P->mx_sweeps; sw++) {
        MxResidArgs r2 = r;
        r2.mode = MX_RES_TRIAL2;
        r2.G2 = P->Gs;
        launch_mx_resid(r2, L, s);
        mx_apply(P, MX_OP_AINV, 0, L, MX_EPI_ADD, P->Gs, P->Gs);
    }
    P->nl1 = P->nl2 = out;  // state.nonlin[2] .= state.nonlin[1]
    mx_vort2flux(P, P->Gs, P->qs1, true);
}

// advance!(state, prob, t) -- timestepping.jl:35-67.  hist: optional CD/CL slot.
static void step_body(ibpm_prob* P, const ibpm_step_scalars& sc, double* cdh, double* clh)
{
    hipStream_t s = P->stream;
    const GridDev& g = P->g;
    const long nG = g.nG;
    const int L = P->mg;
    if (P->motion == IBPM_MOTION_MOVING_GRID) {
        // base_flux! MovingGrid (timestepping.jl:284-298)
        double Om = -sc.theta_dot;
        double al = -sc.theta;
        double Ux0 = sc.U * std::cos(al) - sc.V * std::sin(al);
        double Uy0 = sc.U * std::sin(al) + sc.V * std::cos(al);
        Ux0 += sc.Ux * std::cos(sc.inclination) - sc.Uy * std::sin(sc.inclination);
        Uy0 += sc.Ux * std::sin(sc.inclination) + sc.Uy * std::cos(sc.inclination);
        BaseFluxArgs b{};
        b.q0 = P->q0; b.mg = L; b.Ux0 = Ux0; b.Uy0 = Uy0; b.Om = Om; b.offx = P->offx; b.offy = P->offy; b.g = g;
        launch_base_flux_moving(b, s);
        P->q0_uniform = false;
    }
    const bool mixed = P->prec >= 0;
    if (mixed) step_mixed(P);
    else trial_dev(P);
    // ---- boundary_forces! (timestepping.jl:154-163) and update_stress!
    if (P->nbt > 0) {
        GatherArgs ga{};
        ga.q = P->qs1; ga.q0 = P->q0; ga.w = P->w; ga.bidx = P->bidx; ga.out = P->Ftmp; ga.nbt = P->nbt; ga.g = g;
        if (!mixed) trial_flux_from_psi(P, ga);
        launch_gather(ga, s);
        ForceArgs fa{};
        fa.Fb = P->Fb; fa.ds = P->ds_d; fa.boff = P->boff_d; fa.fb = P->fbv; fa.CD = P->CD; fa.CL = P->CL;
        fa.cd_hist = cdh; fa.cl_hist = clh; fa.nbodies = P->nbodies; fa.nbt = P->nbt; fa.hdt = g.h / P->dt;
        if (cdh) {  // history entry taken from the device slot
            fa.slot = P->slot;
            fa.nbh = std::max(P->nbodies, 1);
        }
        // ---- project_circ! (timestepping.jl:204-223): Γ = Γs; Γ[:,1] -= Ainv rot(reg(F̃b));
        // the FP64 path forms update_stress! in the launch of reg(F̃b)
        if (mixed) {
            launch_gemv_forces(P->BinvT, P->Ftmp, P->Fb, P->nf, fa, s);
            mx_project_forces(P, P->Fb, P->Gs, true);
        } else {
            launch_gemv(P->BinvT, P->Ftmp, P->Fb, P->nf, s);
            project_forces(P, P->Fb, P->Gs, true, &fa);
        }
    }
    std::swap(P->G, P->Gs);
    // ---- vort2flux!(state.ψ, state.q, state.Γ) (timestepping.jl:64); the
    // FP64 path leaves q to be formed from ψ by its readers
    if (mixed) {
        mx_vort2flux(P, P->G, P->q, false);
        P->q_stale = false;
    } else {
        // Γ = Γs - X(Ŷ) on level 1 (Ŷ: the projection's y-pass output, in T1's
        // level 1) and Γs = X(T2) off coarsify!'s rows, as in the trial chain.
        // ψ differs from a full vort2flux!(Γ) by rounding: checkpoints keep ψ.
        vort2flux_dev(P, P->G, P->psi, P->q, CURL_NONE, P->xreuse ? P->T2 : nullptr,
                      P->xreuse && P->nbt > 0 ? P->T1 : nullptr);
    }
    (void)nG;
}

// get_trial_state! (timestepping.jl:86-129), all levels at once: Γs, the
// AB2 history shift and the trial ψ in state.ψ, of which only the lines E and
// the vort2flux! chain read are stored (the step's own vort2flux! rewrites
// every line).  The trial flux qs[:,1] is not stored: E forms it from that ψ
// (trial_flux_from_psi).
static void trial_dev(ibpm_prob* P)
{
    hipStream_t s = P->stream;
    const GridDev& g = P->g;
    const int L = P->mg;
    const int out = (P->nl1 != P->nl2) ? P->nl1 : 1 - P->nl2;
    NLArgs na{};
    na.q = P->q; na.q0 = P->q0; na.G = P->G; na.out = P->nl[out];
    nl_flux_source(P, na);
    na.nlev = L; na.lev0 = 0; na.mg = L; na.g = g;
    set_nl_q0(P, na);
    launch_nonlinear(na, s);
    // rhs = A Γ + ... (A by its 5-point stencil) and Ainv's first x-transform in one pass
    XArgs t = xargs(P);
    t.src = P->G; t.dst = P->T2; t.nlev = L; t.lev0 = 0;
    t.n1 = P->nl[out]; t.n0 = P->nl[P->nl2]; t.gam = P->G;
    t.c1 = P->beta[0] * P->dt;
    t.c2 = P->beta[1] * P->dt;
    t.dt = P->dt;
    t.Re = P->Re;
    launch_xpass(X_TRIAL, t, s);
    timed_ypass(P, yargs(P, Y_AINV, 0, L, P->T2), P->ytile, 1);
    XArgs c = xargs(P);
    c.src = P->T2; c.dst = P->Gs; c.nlev = L; c.lev0 = 0;
    launch_xpass(X_PLAIN, c, s);
    P->nl1 = P->nl2 = out;  // state.nonlin[2] .= state.nonlin[1]
    // Γs = X(T2) on every line coarsify! leaves alone: the chain x-transforms
    // only the centre rows, and reads the rest as 2 nx T2
    vort2flux_dev(P, P->Gs, P->psi, P->qs1, CURL_TRIAL, P->xreuse ? P->T2 : nullptr);
}

// E of the trial flux qs[:,1] = curl!(ψ of the trial) read from ψ itself
// (levels 1 and 2), as trial_dev leaves it
static void trial_flux_from_psi(ibpm_prob* P, GatherArgs& ga)
{
    ga.q = nullptr;
    ga.psi = P->psi;
    ga.psic = P->mg > 1 ? P->psi + P->g.nG : nullptr;
}

// move_body!(::Type{MotionFunction}, body, t) for every body (move-body-utils.jl:
// 123-133, with get_transformation :73-97); kin = {x, y, θ, ẋ, ẏ, θ̇} per body:
//   xb = pos + Rx x0, Rx = [cθ -sθ; sθ cθ];  ub = vel + Ru x0, Ru = θ̇ [-sθ -cθ; cθ -sθ]
static void move_bodies(ibpm_prob* P, const double* kin)
{
    const int nbt = P->nbt;
    for (int b = 0; b < P->nbodies; b++) {
        const double* k = kin + 6L * b;
        const double c = std::cos(k[2]), sn = std::sin(k[2]);
        const double Rx11 = c, Rx12 = -sn, Rx21 = sn, Rx22 = c;
        const double Ru11 = k[5] * -sn, Ru12 = k[5] * -c, Ru21 = k[5] * c, Ru22 = k[5] * -sn;
        for (int i = P->boff[b]; i < P->boff[b + 1]; i++) {
            const double x0 = P->x0[i], y0 = P->x0[nbt + i];
            P->xb[i] = k[0] + (Rx11 * x0 + Rx12 * y0);
            P->xb[nbt + i] = k[1] + (Rx21 * x0 + Rx22 * y0);
            P->ub[i] = k[3] + (Ru11 * x0 + Ru12 * y0);
            P->ub[nbt + i] = k[4] + (Ru21 * x0 + Ru22 * y0);
        }
    }
}

// update_bodies! for MotionFunction (timestepping.jl:386-398): move the bodies,
// then E = setup_reg on the device.  The E' support moved, so the scratch that
// the scatter and rot! leave untouched outside it is cleared again.
static int update_bodies_dev(ibpm_prob* P, const double* kin)
{
    const std::vector<double> xb_prev = P->xb, ub_prev = P->ub;
    move_bodies(P, kin);
    if (!setup_reg_host(P)) {
        // keep the bodies and E of the last good position
        P->xb = xb_prev;
        P->ub = ub_prev;
        (void)setup_reg_host(P);
        return fail(P->ctx, IBPM_E_ARG, "a moving body came within 6 cells of the domain edge");
    }
    HIPTRY(P->ctx, upload_reg(P));
    launch_fill(P->qreg, P->g.nq, 0.0, P->stream);
    launch_fill(P->Gtmp, P->g.nG, 0.0, P->stream);
    return IBPM_OK;
}

// B = [B_times!(e_1) ... B_times!(e_nf)] (laplacian.jl:143-169) for the
// current E, in batches of bk unit forces; every field of a batch goes through
// the operators of one B_times!: E' e_f, rot!, Ainv on level 1, vort2flux!
// (coarsify chain, Δinv coarse to fine with the ψ boundary values, curl on
// level 1) and E.  The result lands in P->Bh.
static int assemble_b(ibpm_prob* P)
{
    hipStream_t s = P->stream;
    const GridDev& g = P->g;
    const long nG = g.nG, nq = g.nq;
    const int L = P->mg, nf = P->nf, K = P->bk;
    for (int f0 = 0; f0 < nf; f0 += K) {
        const int kc = std::min(K, nf - f0);
        launch_fill(P->bq, nq * kc, 0.0, s);
        launch_unit_reg(P->bq, nq, kc, f0, P->bidx, P->w, P->nbt, g, s);
        launch_fill(P->bgt, nG * kc, 0.0, s);  // rot! writes the E' support box only
        launch_rot_batch(P->bq, nq, P->bgt, nG, kc, P->rb_i0, P->rb_i1, P->rb_j0, P->rb_j1, g, s);
        // Γ[:, 1] = Ainv rot(E' e_f); the rows outside the box are zero
        XArgs a = xargs(P);
        a.src = P->bgt; a.dst = P->bt1; a.nlev = kc; a.lev0 = 0; a.fixlev = 1;
        a.zlo = P->rb_j0 - 1; a.zhi = std::min(P->rb_j1 - 1, g.n2);
        launch_xpass(X_PLAIN, a, s);
        YArgs y = yargs(P, Y_AINV, 0, kc, P->bt1);
        y.fixlev = 1;
        launch_ypass(y, P->ytile, s);
        XArgs c = xargs(P);
        c.src = P->bt1; c.dst = P->bG; c.nlev = kc; c.lev0 = 0; c.fixlev = 1;
        launch_xpass(X_PLAIN, c, s);
        // vort2flux!: coarsify into the centres of the coarse levels (their
        // rings stay zero), then Δinv from the coarsest level down
        for (int lev = 2; lev <= L; lev++) {
            CoarsifyArgs cr{};
            const CoarsifyRange& r = P->crange[0];
            cr.Gc = P->bG + (long)(lev - 1) * K * nG;
            cr.G = P->bG + (long)(lev - 2) * K * nG;
            cr.i0 = r.i0; cr.ic0 = r.ic0; cr.ci = r.ci; cr.j0 = r.j0; cr.jc0 = r.jc0; cr.cj = r.cj;
            cr.nb = kc; cr.bstride = nG; cr.g = g;
            launch_coarsify(cr, s);
        }
        for (int l = L - 1; l >= 0; l--) {
            XArgs a1 = xargs(P);
            a1.src = P->bG + (long)l * K * nG; a1.dst = P->bt1; a1.nlev = kc; a1.lev0 = l; a1.fixlev = 1;
            launch_xpass(X_PLAIN, a1, s);
            YArgs y1 = yargs(P, Y_DINV, l, kc, P->bt1);
            y1.fixlev = 1;
            if (l < L - 1) {  // apply_bc! from ψ of level l + 1, in the y-pass
                y1.bcr = P->bpsi + (long)(l + 1) * K * nG;
                y1.bcstride = nG;
                y1.edge = P->bedge;
                y1.edstride = 2L * g.n2;
            }
            launch_ypass(y1, P->ytile, s);
            XArgs c1 = xargs(P);
            c1.src = P->bt1; c1.dst = P->bpsi + (long)l * K * nG; c1.nlev = kc; c1.lev0 = l; c1.fixlev = 1;
            if (l > 0) {
                c1.edge = P->bedge;
                c1.edstride = 2L * g.n2;
            }
            launch_xpass(X_PLAIN, c1, s);
        }
        CurlArgs ca{};
        ca.psi = P->bpsi; ca.q = P->bq; ca.qstride = nq; ca.nlev = kc; ca.lev0 = 0; ca.mg = L;
        ca.colbatch = 1; ca.psl = (long)K * nG; ca.psb = nG; ca.g = g;
        launch_curl(ca, s);
        GatherArgs ga{};
        ga.q = P->bq; ga.q0 = nullptr; ga.w = P->w; ga.bidx = P->bidx; ga.out = P->Bd + (long)f0 * nf;
        ga.nbt = P->nbt; ga.nb = kc; ga.qstride = nq; ga.ostride = nf; ga.g = g;
        launch_gather(ga, s);
    }
    HIPTRY(P->ctx, hipGetLastError());
    HIPTRY(P->ctx, hipMemcpyAsync(P->Bh.data(), P->Bd, sizeof(double) * P->Bh.size(), hipMemcpyDeviceToHost, s));
    HIPTRY(P->ctx, hipStreamSynchronize(s));
    return IBPM_OK;
}

// c = B u with the assembled B (column-major)
static void b_matvec(const std::vector<double>& B, int nf, const double* u, double* c)
{
    std::fill(c, c + nf, 0.0);
    for (int j = 0; j < nf; j++) {
        const double uj = u[j];
        const double* col = B.data() + (size_t)j * nf;
        for (int i = 0; i < nf; i++) c[i] += col[i] * uj;
    }
}

static double dot_n(const double* a, const double* b, int n)
{
    double acc = 0.0;
    for (int k = 0; k < n; k++) acc += a[k] * b[k];
    return acc;
}

// advance! for MotionFunction bodies (timestepping.jl:35-67): update_bodies!,
// get_trial_state!, boundary_forces! for V <: Motion (timestepping.jl:176-186:
// F̃work = E(qs + q0) - ub h, F̃b = Binv F̃work with Binv the CG map of
// laplacian.jl:137-138), update_stress!, project_circ!, vort2flux!.  The CG
// recurrences of cg! (IterativeSolvers.jl 0.9.1: initial guess = the state's
// F̃b, tolerance max(1e-12 |b|, 0), maxiter 5000) run on the host over the
// 2 nb_total-long vectors; their products B u use B as assembled on the
// device for this step's E (one batched pass of all unit forces through
// B_times!'s operators), which equals the reference's matrix-free product
// up to rounding.  Synchronous.
static int step_mf(ibpm_prob* P, const ibpm_step_scalars& sc, double* cdh, double* clh)
{
    ibpm_ctx* C = P->ctx;
    hipStream_t s = P->stream;
    const GridDev& g = P->g;
    const int nf = P->nf;
    if (!sc.body_kin) return fail(C, IBPM_E_ARG, "MotionFunction steps need body_kin (6 values per body)");
    int rc = update_bodies_dev(P, sc.body_kin);
    if (rc != IBPM_OK) return rc;
    // q formed first, so that a step that fails below leaves the state's q
    ensure_q(P);
    trial_dev(P);
    if (nf > 0) {
        // b = E(qs + q0) - ub h, and the initial guess: the state's F̃b
        GatherArgs ga{};
        ga.q = P->qs1; ga.q0 = P->q0; ga.w = P->w; ga.bidx = P->bidx; ga.out = P->Fw; ga.nbt = P->nbt; ga.g = g;
        trial_flux_from_psi(P, ga);
        launch_gather(ga, s);
        HIPTRY(C, hipMemcpyAsync(P->pin + nf, P->Fw, sizeof(double) * nf, hipMemcpyDeviceToHost, s));
        HIPTRY(C, hipMemcpyAsync(P->pin, P->Fb, sizeof(double) * nf, hipMemcpyDeviceToHost, s));
        rc = assemble_b(P);  // synchronises the stream
        if (rc != IBPM_OK) return rc;
        std::vector<double> b(nf), x(nf), r(nf), u(nf, 0.0), c(nf);
        for (int k = 0; k < nf; k++) b[k] = P->pin[nf + k] - P->ub[k] * g.h;
        std::copy(P->pin, P->pin + nf, x.begin());
        // cg_iterator! with initially_zero = false: r = b - B x
        b_matvec(P->Bh, nf, x.data(), c.data());
        for (int k = 0; k < nf; k++) r[k] = b[k] - c[k];
        double res = std::sqrt(dot_n(r.data(), r.data(), nf));
        const double tol = std::max(1e-12 * std::sqrt(dot_n(b.data(), b.data(), nf)), 0.0);
        double prev = 1.0;
        int it = 0;
        // a non-finite residual ends the solve (the reference would run all
        // maxiter products to the same non-finite F̃b)
        while (!(it >= 5000 || res <= tol) && std::isfinite(res)) {
            const double beta = (res * res) / (prev * prev);
            for (int k = 0; k < nf; k++) u[k] = r[k] + beta * u[k];
            b_matvec(P->Bh, nf, u.data(), c.data());
            const double alpha = (res * res) / dot_n(u.data(), c.data(), nf);
            for (int k = 0; k < nf; k++) x[k] += alpha * u[k];
            for (int k = 0; k < nf; k++) r[k] -= alpha * c[k];
            prev = res;
            res = std::sqrt(dot_n(r.data(), r.data(), nf));
            it++;
        }
        P->cg_iters = it;
        P->cg_res = res;
        P->cg_tol = tol;
        std::copy(x.begin(), x.end(), P->pin);
        HIPTRY(C, hipMemcpyAsync(P->Fb, P->pin, sizeof(double) * nf, hipMemcpyHostToDevice, s));
        // ---- update_stress!
        ForceArgs fa{};
        fa.Fb = P->Fb; fa.ds = P->ds_d; fa.boff = P->boff_d; fa.fb = P->fbv; fa.CD = P->CD; fa.CL = P->CL;
        fa.cd_hist = cdh; fa.cl_hist = clh; fa.nbodies = P->nbodies; fa.nbt = P->nbt; fa.hdt = g.h / P->dt;
        if (cdh) {
            fa.slot = P->slot;
            fa.nbh = std::max(P->nbodies, 1);
        }
        launch_forces(fa, s);
        // ---- project_circ!
        project_forces(P, P->Fb, P->Gs, true);
    }
    std::swap(P->G, P->Gs);
    vort2flux_dev(P, P->G, P->psi, P->q, CURL_NONE);
    HIPTRY(C, hipGetLastError());
    // the pinned staging is reused by the next step: drain before returning
    HIPTRY(C, hipStreamSynchronize(s));
    return IBPM_OK;
}

static void graphs_clear(ibpm_prob* P)
{
    for (auto& e : P->gcache) (void)hipGraphExecDestroy(e.exec);
    P->gcache.clear();
}

// One advance! step on the problem's stream.  cdh/clh: bases of the CD/CL
// histories (entry taken from the device slot), or null.  Static bodies are
// replayed from a captured graph of the step when one exists for the
// current buffer roles, else the step is captured (and then run) once.
static void step_dev(ibpm_prob* P, const ibpm_step_scalars& sc, double* cdh, double* clh)
{
    const bool can_graph = P->graphs && P->motion == IBPM_MOTION_STATIC && !P->prof && cdh;
    if (!can_graph) {
        step_body(P, sc, cdh, clh);
        return;
    }
    const int out = (P->nl1 != P->nl2) ? P->nl1 : 1 - P->nl2;
    for (auto& e : P->gcache) {
        if (e.G == P->G && e.out == out && e.nl2 == P->nl2 && e.qs == P->q_stale) {
            if (hipGraphLaunch(e.exec, P->stream) != hipSuccess) P->async_err = true;
            // the host side of step_body: nonlin roles, Γ <-> Γs, q left to ψ (FP64)
            P->nl1 = P->nl2 = out;
            std::swap(P->G, P->Gs);
            P->q_stale = P->prec < 0;
            return;
        }
    }
    ibpm_prob::StepGraph e{P->G, out, P->nl2, P->q_stale, nullptr};
    double* G0 = P->G;
    double* Gs0 = P->Gs;
    const int nl1 = P->nl1, nl2 = P->nl2;
    const bool qs0 = P->q_stale;
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        step_body(P, sc, cdh, clh);
        ok = hipStreamEndCapture(P->stream, &graph) == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (ok) ok = hipGraphLaunch(e.exec, P->stream) == hipSuccess;
    if (ok) {
        P->gcache.push_back(e);
        return;
    }
    // capture unavailable: undo the host side of the captured body and run it directly
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    (void)hipGetLastError();
    P->graphs = false;
    P->G = G0; P->Gs = Gs0; P->nl1 = nl1; P->nl2 = nl2; P->q_stale = qs0;
    step_body(P, sc, cdh, clh);
}

// ---------------------------------------------------------------------------
// Tangent-linear and adjoint advance! (MovingGrid bodies, FP64).
//
// A step maps the state (Γ, N0 = nonlin[2]) and the step's scalars s_k to
// (Γ', N1), with q = vort2flux!(Γ) taken inside it (the state's q is always
// that function of the state's Γ).  Everything but the product of the
// nonlinear term (bilinear in q + q0 and Γ with its boundary ring) and the
// base flux q0(s_k) (timestepping.jl:284-317) is linear, so the tangent and
// the adjoint of a step need only its input state and scalars:
//   δq = V_q(δΓ), δq0 = (∂q0/∂s) δs,
//   δN1 = rot(D(avg(δq + δq0), Γ) + D(avg(q + q0), δΓ)) / (4 hc^2),
//   δΓs = Ainv(A δΓ + β1 dt δN1 + β2 dt δN0 + rhsbc(δΓ)),
//   δF̃b = Binv E(V_q1(δΓs) + δq0), δCD = 2 (h/dt) Σ δF̃b_x, δCL likewise,
//   δΓ' = C(C(δΓs) - Ainv rot E' δF̃b)    (C: coarsify!, V: vort2flux!),
// and the adjoint step applies the transposes in reverse order.  The
// objective is J = Σ_k (wcd_k CD_k + wcl_k CL_k) over the steps.
// ---------------------------------------------------------------------------
template <typename T>
static hipError_t walloc(AdjWork* W, T** p, size_t n)
{
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (n > 0 ? n : 1));
    if (e == hipSuccess) {
        W->allocs.push_back(v);
        *p = (T*)v;
    }
    return e;
}

static void adj_free(ibpm_prob* P)
{
    if (!P->adj) return;
    for (void* v : P->adj->allocs) (void)hipFree(v);
    delete P->adj;
    P->adj = nullptr;
}

static int adj_setup(ibpm_prob* P, int nsteps, int nsnap)
{
    const GridDev& g = P->g;
    const long nG = g.nG, nq = g.nq, L = P->mg;
    if (!P->adj) {
        AdjWork* W = new AdjWork();
        P->adj = W;
#define WA(ptr, n) HIPTRY(P->ctx, walloc(W, &(ptr), (size_t)(n)))
        WA(W->dG, nG * L); WA(W->dGs, nG * L); WA(W->dN[0], nG * L); WA(W->dN[1], nG * L); WA(W->tG, nG * L);
        WA(W->dpsi, nG * L); WA(W->dq, nq * L); WA(W->dq0, nq * L); WA(W->dqs1, nq); WA(W->Q, nq); WA(W->dQ, nq);
        WA(W->fq, nq); WA(W->F1, P->nf); WA(W->F2, P->nf); WA(W->BinvC, (size_t)P->nf * P->nf);
        WA(W->bc, g.nbc); WA(W->bcb, g.nbc); WA(W->tG1, nG);
        WA(W->Gb, nG * L); WA(W->Gbo, nG * L); WA(W->Nb1, nG * L); WA(W->Nb0, nG * L); WA(W->Rb, nG * L);
        WA(W->psib, nG * L); WA(W->qb, nq * L); WA(W->q0b, nq * L); WA(W->qs1b, nq);
        WA(W->part, 3L * adj_q0_reduce_parts(g, P->mg));
        WA(W->aux, nq * L + P->nf + 2L * P->nbodies);
#undef WA
        // Binv' as row-major rows: the column-major Binv itself
        HIPTRY(P->ctx, hipMemcpy(W->BinvC, P->binv.data(), sizeof(double) * P->binv.size(), hipMemcpyHostToDevice));
        launch_fill(W->qs1b, nq, 0.0, P->stream);  // E' writes only its support cells
    }
    AdjWork* W = P->adj;
    if (nsteps > W->nsteps_cap) {
        HIPTRY(P->ctx, walloc(W, &W->sums, 3 * (size_t)nsteps));
        W->nsteps_cap = nsteps;
    }
    while ((int)W->snap.size() < nsnap) {
        double* sp = nullptr;
        HIPTRY(P->ctx, walloc(W, &sp, 3 * (size_t)(nG * L)));
        W->snap.push_back(sp);
    }
    return IBPM_OK;
}

// base-flux parameters of a MovingGrid step (timestepping.jl:284-298)
struct MGFlux {
    double Ux0, Uy0, Om;
};
static MGFlux mg_flux(const ibpm_step_scalars& sc)
{
    const double Om = -sc.theta_dot, al = -sc.theta;
    double Ux0 = sc.U * std::cos(al) - sc.V * std::sin(al);
    double Uy0 = sc.U * std::sin(al) + sc.V * std::cos(al);
    Ux0 += sc.Ux * std::cos(sc.inclination) - sc.Uy * std::sin(sc.inclination);
    Uy0 += sc.Ux * std::sin(sc.inclination) + sc.Uy * std::cos(sc.inclination);
    return {Ux0, Uy0, Om};
}
// d(Ux0, Uy0, Ω) for perturbations d[0..6] = {Ux, Uy, inclination, U, V, θ, θ̇}
static MGFlux mg_flux_tangent(const ibpm_step_scalars& sc, const double* d)
{
    const double al = -sc.theta, ca = std::cos(al), sa = std::sin(al);
    const double cf = std::cos(sc.inclination), sf = std::sin(sc.inclination);
    const double dal = -d[5];
    MGFlux r;
    r.Ux0 = d[3] * ca - d[4] * sa + (-sc.U * sa - sc.V * ca) * dal + d[0] * cf - d[1] * sf +
            (-sc.Ux * sf - sc.Uy * cf) * d[2];
    r.Uy0 = d[3] * sa + d[4] * ca + (sc.U * ca - sc.V * sa) * dal + d[0] * sf + d[1] * cf +
            (sc.Ux * cf - sc.Uy * sf) * d[2];
    r.Om = -d[6];
    return r;
}
// the transpose: cotangents of the 7 scalars from (Ūx0, Ūy0, Ω̄)
static void mg_flux_adjoint(const ibpm_step_scalars& sc, const double* b, double* out)
{
    const double al = -sc.theta, ca = std::cos(al), sa = std::sin(al);
    const double cf = std::cos(sc.inclination), sf = std::sin(sc.inclination);
    const double ux = b[0], uy = b[1], om = b[2];
    out[0] = ux * cf + uy * sf;
    out[1] = -ux * sf + uy * cf;
    out[2] = ux * (-sc.Ux * sf - sc.Uy * cf) + uy * (sc.Ux * cf - sc.Uy * sf);
    out[3] = ux * ca + uy * sa;
    out[4] = -ux * sa + uy * ca;
    out[5] = -(ux * (-sc.U * sa - sc.V * ca) + uy * (sc.U * ca - sc.V * sa));
    out[6] = -om;
}

static void set_q0(ibpm_prob* P, double* q0, const MGFlux& f)
{
    BaseFluxArgs b{};
    b.q0 = q0; b.mg = P->mg; b.Ux0 = f.Ux0; b.Uy0 = f.Uy0; b.Om = f.Om; b.offx = P->offx; b.offy = P->offy; b.g = P->g;
    launch_base_flux_moving(b, P->stream);
}

// x = L b over every level (batched passes), L one of Δinv / A / Ainv
static void lap_all(ibpm_prob* P, int op, const double* b, double* x)
{
    hipStream_t s = P->stream;
    XArgs a = xargs(P);
    a.src = b; a.dst = P->T1; a.nlev = P->mg; a.lev0 = 0;
    launch_xpass(X_PLAIN, a, s);
    launch_ypass(yargs(P, op, 0, P->mg, P->T1), P->ytile, s);
    XArgs c = xargs(P);
    c.src = P->T1; c.dst = x; c.nlev = P->mg; c.lev0 = 0;
    launch_xpass(X_PLAIN, c, s);
}

static CoarsifyArgs coarsify_args(ibpm_prob* P, double* Gb, int lev)  // lev: the coarse level, 2..L (1-based)
{
    CoarsifyArgs c{};
    const CoarsifyRange& r = P->crange[0];
    c.Gc = Gb + (long)(lev - 1) * P->g.nG;
    c.G = Gb + (long)(lev - 2) * P->g.nG;
    c.i0 = r.i0; c.ic0 = r.ic0; c.ci = r.ci; c.j0 = r.j0; c.jc0 = r.jc0; c.cj = r.cj;
    c.g = P->g;
    return c;
}

static void coarsify_all(ibpm_prob* P, double* Gb)
{
    for (int lev = 2; lev <= P->mg; lev++) launch_coarsify(coarsify_args(P, Gb, lev), P->stream);
}

static void coarsify_all_T(ibpm_prob* P, double* Gb)
{
    for (int lev = P->mg; lev >= 2; lev--) {
        CoarsifyArgs c = coarsify_args(P, Gb, lev);
        launch_adj_coarsify_T(Gb + (long)(lev - 2) * P->g.nG, c.Gc, c, P->stream);
    }
}

// chainᵀ of vort2flux! (curl.jl:141-162): ḡ = Δinv ψ̄ level by level, fine
// to coarse, ψ̄ of the coarser level += get_bc!ᵀ apply_bc!ᵀ ḡ.  psib is consumed.
static void chain_T(ibpm_prob* P, double* psib, double* gout)
{
    const long nG = P->g.nG;
    const int L = P->mg;
    AdjWork* W = P->adj;
    for (int l = 0; l < L; l++) {
        lap_level(P, Y_DINV, l, psib + l * nG, gout + l * nG, false);
        if (l < L - 1) {
            launch_adj_apply_bc_T(gout + l * nG, 1.0, W->bcb, P->g, P->stream);
            launch_adj_get_bc_T(W->bcb, 1.0, psib + (l + 1) * nG, P->g, P->stream);
        }
    }
}

// ψ̄ from q̄ through the multigrid curl! of levels 0 .. nl-1 (q̄ level l at qb + l nq)
static void curl_T(ibpm_prob* P, const double* qb, int nl, double* psib)
{
    const long nG = P->g.nG, nq = P->g.nq;
    AdjWork* W = P->adj;
    launch_fill(psib, nG * P->mg, 0.0, P->stream);
    for (int l = 0; l < nl; l++) {
        launch_adj_rot(qb + l * nq, psib + l * nG, 1.0, true, P->g, P->stream);
        if (l < P->mg - 1) {
            launch_adj_curl_bc_T(qb + l * nq, W->bcb, P->g, P->stream);
            launch_adj_get_bc_T(W->bcb, 1.0, psib + (l + 1) * nG, P->g, P->stream);
        }
    }
}

// the nonlinear term of every level, linearised about the state (q, q0, Γ):
// dN = rot(D(avg(dq + dq0), Γ) + D(avg(q + q0), dG)) / (4 hc^2)
static void nonlinear_tangent(ibpm_prob* P, const double* dG, const double* dq, const double* dq0, double* dN)
{
    const GridDev& g = P->g;
    const long nG = g.nG, nq = g.nq;
    const int L = P->mg;
    AdjWork* W = P->adj;
    hipStream_t s = P->stream;
    for (int l = 0; l < L; l++) {
        const double* Gl = P->G + l * nG;
        const double* Gc = l < L - 1 ? Gl + nG : nullptr;
        const double* dGl = dG + l * nG;
        const double* dGc = l < L - 1 ? dGl + nG : nullptr;
        launch_adj_avg_flux(P->q + l * nq, P->q0 + l * nq, W->Q, g, s);
        launch_adj_avg_flux(dq + l * nq, dq0 + l * nq, W->dQ, g, s);
        launch_adj_dprod(W->dQ, Gl, Gc, W->fq, false, g, s);
        launch_adj_dprod(W->Q, dGl, dGc, W->fq, true, g, s);
        const double hc = level_hc(g.h, l);
        launch_adj_rot(W->fq, dN + l * nG, 1.0 / (4.0 * (hc * hc)), false, g, s);
    }
}

// One tangent-linear step about the working state (the input of step k).
// W->dG / W->dN[n0] hold (δΓ, δN0); afterwards W->dG holds δΓ' and the new
// δN0 is W->dN[1 - n0].  Σ δF̃b_x, Σ δF̃b_y go to sums[2k], sums[2k+1].
static void tlm_step(ibpm_prob* P, const ibpm_step_scalars& sc, const double* ds, int k, int n0)
{
    const GridDev& g = P->g;
    const long nG = g.nG;
    const int L = P->mg;
    AdjWork* W = P->adj;
    hipStream_t s = P->stream;
    ensure_q(P);  // the working state's q (a forward step leaves it to ψ)
    set_q0(P, P->q0, mg_flux(sc));
    set_q0(P, W->dq0, mg_flux_tangent(sc, ds));
    // δq = V_q(δΓ) on a copy (vort2flux! coarsifies its input)
    launch_copy(W->tG, W->dG, nG * L, s);
    vort2flux_dev(P, W->tG, W->dpsi, W->dq, CURL_ALL);
    double* dN1 = W->dN[1 - n0];
    nonlinear_tangent(P, W->dG, W->dq, W->dq0, dN1);
    // δΓs = Ainv(A δΓ + β1 dt δN1 + β2 dt δN0 + rhsbc(δΓ)): the trial passes
    XArgs t = xargs(P);
    t.src = W->dG; t.dst = P->T2; t.nlev = L; t.lev0 = 0;
    t.n1 = dN1; t.n0 = W->dN[n0]; t.gam = W->dG;
    t.c1 = P->beta[0] * P->dt; t.c2 = P->beta[1] * P->dt; t.dt = P->dt; t.Re = P->Re;
    launch_xpass(X_TRIAL, t, s);
    launch_ypass(yargs(P, Y_AINV, 0, L, P->T2), P->ytile, s);
    XArgs c = xargs(P);
    c.src = P->T2; c.dst = W->dGs; c.nlev = L; c.lev0 = 0;
    launch_xpass(X_PLAIN, c, s);
    vort2flux_dev(P, W->dGs, W->dpsi, W->dqs1, CURL_LEVEL1);
    // δF̃b = Binv E(δqs1 + δq0)
    GatherArgs ga{};
    ga.q = W->dqs1; ga.q0 = W->dq0; ga.w = P->w; ga.bidx = P->bidx; ga.out = W->F1; ga.nbt = P->nbt; ga.g = g;
    launch_gather(ga, s);
    launch_adj_gemv(P->BinvT, W->F1, W->F2, P->nf, s);
    launch_adj_sum_halves(W->F2, P->nf, W->sums + 2L * k, s);
    project_forces(P, W->F2, W->dGs, true);
    std::swap(W->dG, W->dGs);
    coarsify_all(P, W->dG);
}

// One adjoint step about the working state (the input of step k): from
// (Γ̄ of its output Γ' in W->Gb, N̄ of its N1 in W->Nb1) to those of its
// input (W->Gb, W->Nb1 on return); Ūx0, Ūy0, Ω̄ go to sums[3k..3k+2].
static void adj_step(ibpm_prob* P, const ibpm_step_scalars& sc, double wcd, double wcl, int k)
{
    const GridDev& g = P->g;
    const long nG = g.nG, nq = g.nq;
    const int L = P->mg;
    AdjWork* W = P->adj;
    hipStream_t s = P->stream;
    ensure_q(P);  // the working state's q (a forward step leaves it to ψ)
    set_q0(P, P->q0, mg_flux(sc));
    // Γ' = C(Γs) - Ainv rot E' F (level 1), then the state's coarsify!
    coarsify_all_T(P, W->Gb);
    // F̄ = -E curl(Ainv Γ̄'_1) + the objective's weights
    lap_level(P, Y_AINV, 0, W->Gb, W->tG1, false);
    launch_adj_curl_sg(W->tG1, W->dqs1, -1.0, g, s);
    GatherArgs ga{};
    ga.q = W->dqs1; ga.q0 = nullptr; ga.w = P->w; ga.bidx = P->bidx; ga.out = W->F1; ga.nbt = P->nbt; ga.g = g;
    launch_gather(ga, s);
    const double cD = 2.0 * g.h / P->dt;
    launch_adj_add_halves(W->F1, P->nf, cD * wcd, cD * wcl, s);
    // ū = Binv' F̄; q̄s1 = E' ū (also q̄0 of level 1)
    launch_adj_gemv(W->BinvC, W->F1, W->F2, P->nf, s);
    ScatterArgs sa{};
    sa.cell = P->sc_cell; sa.off = P->sc_off; sa.ent_f = P->sc_f; sa.ent_w = P->sc_w;
    sa.fb = W->F2; sa.q = W->qs1b; sa.ncell = P->sc_ncell;
    launch_scatter(sa, s);
    launch_fill(W->q0b, nq * L, 0.0, s);
    launch_copy(W->q0b, W->qs1b, nq, s);
    // Γ̄s = C'(Γ̄' + chain' curl_1' q̄s1)
    curl_T(P, W->qs1b, 1, W->psib);
    chain_T(P, W->psib, W->tG);
    launch_adj_axpy(W->tG, W->Gb, 1.0, nG * L, s);
    coarsify_all_T(P, W->tG);
    // R̄ = Ainv Γ̄s; Γ̄ = A R̄ + rhsbc'; N̄1 += β1 dt R̄; N̄0 = β2 dt R̄
    lap_all(P, Y_AINV, W->tG, W->Rb);
    lap_all(P, Y_A, W->Rb, W->Gbo);
    launch_adj_axpy(W->Nb1, W->Rb, P->beta[0] * P->dt, nG * L, s);
    launch_fill(W->Nb0, nG * L, 0.0, s);
    launch_adj_axpy(W->Nb0, W->Rb, P->beta[1] * P->dt, nG * L, s);
    for (int l = 0; l < L - 1; l++) {
        const double hc = level_hc(g.h, l);
        launch_adj_apply_bc_T(W->Rb + l * nG, 1.0, W->bcb, g, s);
        launch_adj_get_bc_T(W->bcb, 0.25 * P->dt / (P->Re * hc * hc), W->Gbo + (l + 1) * nG, g, s);
    }
    // the nonlinear term of every level
    for (int l = 0; l < L; l++) {
        const double hc = level_hc(g.h, l);
        const double* Gl = P->G + l * nG;
        const double* Gc = l < L - 1 ? Gl + nG : nullptr;
        launch_adj_curl_sg(W->Nb1 + l * nG, W->fq, 1.0 / (4.0 * (hc * hc)), g, s);
        launch_adj_avg_flux(P->q + l * nq, P->q0 + l * nq, W->Q, g, s);
        launch_adj_dprod_TQ(W->fq, Gl, Gc, W->dQ, g, s);
        launch_adj_dprod_TG(W->fq, W->Q, W->Gbo + l * nG, W->bc, g, s);
        if (l < L - 1) launch_adj_get_bc_T(W->bc, 0.25, W->Gbo + (l + 1) * nG, g, s);
        launch_adj_avg_flux_T(W->dQ, W->qb + l * nq, false, g, s);
        launch_adj_avg_flux_T(W->dQ, W->q0b + l * nq, true, g, s);
    }
    // q = V_q(Γ): Γ̄ += C' chain' curl' q̄
    curl_T(P, W->qb, L, W->psib);
    chain_T(P, W->psib, W->tG);
    coarsify_all_T(P, W->tG);
    launch_adj_axpy(W->Gbo, W->tG, 1.0, nG * L, s);
    launch_adj_q0_reduce(W->q0b, L, P->offx, P->offy, W->part, W->sums + 3L * k, g, s);
    std::swap(W->Gb, W->Gbo);
    std::swap(W->Nb1, W->Nb0);
}

// checkpoints of the working state: (Γ, nonlin[2]) copied out and back; a
// restore recomputes ψ and q by vort2flux! (coarsify! of a coarsified Γ
// rewrites the same values), which is how the forward step left them
// save (or put back) the parts of the state a run overwrites besides Γ,
// nonlin[2], ψ and q: the base flux q0 and the forces F̃b, CD, CL
static void aux_state(ibpm_prob* P, bool save)
{
    AdjWork* W = P->adj;
    double* a = W->aux;
    auto cp = [&](double* x, long n) {
        if (save) launch_copy(a, x, n, P->stream);
        else launch_copy(x, a, n, P->stream);
        a += n;
    };
    cp(P->q0, P->g.nq * P->mg);
    cp(P->Fb, P->nf);
    cp(P->CD, P->nbodies);
    cp(P->CL, P->nbodies);
    if (save) {
        W->q0_uniform = P->q0_uniform; W->q0vx = P->q0vx; W->q0vy = P->q0vy;
    } else {
        P->q0_uniform = W->q0_uniform; P->q0vx = W->q0vx; P->q0vy = W->q0vy;
    }
}

// A checkpoint keeps ψ beside Γ: the step's vort2flux! forms ψ from the
// trial's transforms (equal to vort2flux!(Γ) up to rounding), so recomputing
// it from Γ would not replay the forward run bit for bit.
static void snap_store(ibpm_prob* P, int slot)
{
    const long n = P->g.nG * P->mg;
    double* sp = P->adj->snap[slot];
    launch_copy(sp, P->G, n, P->stream);
    launch_copy(sp + n, P->nl[P->nl2], n, P->stream);
    launch_copy(sp + 2 * n, P->psi, n, P->stream);
}
static void snap_restore(ibpm_prob* P, int slot)
{
    const long n = P->g.nG * P->mg;
    const double* sp = P->adj->snap[slot];
    P->nl1 = P->nl2;
    launch_copy(P->G, sp, n, P->stream);
    launch_copy(P->nl[P->nl2], sp + n, n, P->stream);
    launch_copy(P->psi, sp + 2 * n, n, P->stream);
    P->q_stale = true;  // formed from ψ where it is read
}

static long binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    long r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

// Reverse steps l .. l+t-1 with the state of step l in snapshot `base` and
// `free_` further snapshots after it (binomial checkpointing: with s free
// snapshots and r the least recomputation count with C(s+r, s) >= t, the
// next snapshot goes t - C(s-1+r, s-1) steps ahead, so that the part after
// it is reversible with s-1 snapshots and the part before with r-1 sweeps)
struct AdjRun {
    ibpm_prob* P;
    const ibpm_step_scalars* s;
    const double *wcd, *wcl;
    long fwd_steps = 0;
};
static void rev(AdjRun& R, int l, int t, int base, int free_)
{
    ibpm_prob* P = R.P;
    if (t <= 0) return;
    if (t == 1) {
        snap_restore(P, base);
        adj_step(P, R.s[l], R.wcd[l], R.wcl[l], l);
        return;
    }
    if (free_ == 0) {
        for (int k = t - 1; k >= 0; k--) {
            snap_restore(P, base);
            for (int j = 0; j < k; j++) step_body(P, R.s[l + j], nullptr, nullptr);
            R.fwd_steps += k;
            adj_step(P, R.s[l + k], R.wcd[l + k], R.wcl[l + k], l + k);
        }
        return;
    }
    int r = 0;
    while (binom(free_ + r, free_) < t) r++;
    const int m = std::max(1, t - (int)binom(free_ - 1 + r, free_ - 1));
    snap_restore(P, base);
    for (int j = 0; j < m; j++) step_body(P, R.s[l + j], nullptr, nullptr);
    R.fwd_steps += m;
    snap_store(P, base + 1);
    rev(R, l + m, t - m, base + 1, free_ - 1);
    rev(R, l, m, base, free_);
}

static int adj_check(ibpm_prob* P)
{
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group");
    if (int rc = check_idle(P)) return rc;
    if (P->motion != IBPM_MOTION_MOVING_GRID || P->nbodies != 1)
        return fail(P->ctx, IBPM_E_UNSUPPORTED, "the tangent and adjoint steps are built for one MovingGrid body");
    if (P->prec >= 0 || P->lin)
        return fail(P->ctx, IBPM_E_UNSUPPORTED, "the tangent and adjoint steps run the FP64 nonlinear problem");
    return IBPM_OK;
}

// get_Binv (laplacian.jl:98-119): B[:, j] = B_times!(e_j) on the device,
// then inv(B) on the host.
static int build_binv(ibpm_prob* P)
{
    ibpm_ctx* C = P->ctx;
    const GridDev& g = P->g;
    const int nf = P->nf;
    const long nG = g.nG, nq = g.nq, L = P->mg;
    // All 2 nb unit forces through B_times! in batches (assemble_b): as many
    // fields per batch as fit in 8 GiB of scratch, allocated for the setup only
    const size_t per = sizeof(double) * ((size_t)nq + 2 * (size_t)nG + 2 * (size_t)L * nG + 2 * (size_t)g.n2);
    const int K = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nf, ((size_t)8 << 30) / per, (size_t)32768}));
    double *bq = nullptr, *bgt = nullptr, *bt1 = nullptr, *bG = nullptr, *bpsi = nullptr, *bedge = nullptr,
           *Bd = nullptr;
    auto release = [&]() {
        for (double* p : {bq, bgt, bt1, bG, bpsi, bedge, Bd}) if (p) (void)hipFree(p);
        P->bq = P->bgt = P->bt1 = P->bG = P->bpsi = P->bedge = P->Bd = nullptr;
        P->bk = 0;
    };
    hipError_t e = hipMalloc(&bq, sizeof(double) * K * nq);
    if (e == hipSuccess) e = hipMalloc(&bgt, sizeof(double) * K * nG);
    if (e == hipSuccess) e = hipMalloc(&bt1, sizeof(double) * K * nG);
    if (e == hipSuccess) e = hipMalloc(&bG, sizeof(double) * K * L * nG);
    if (e == hipSuccess) e = hipMalloc(&bpsi, sizeof(double) * K * L * nG);
    if (e == hipSuccess) e = hipMalloc(&bedge, sizeof(double) * K * 2 * (size_t)g.n2);
    if (e == hipSuccess) e = hipMalloc(&Bd, sizeof(double) * (size_t)nf * nf);
    if (e != hipSuccess) {
        release();
        return fail(C, IBPM_E_HIP, "hipMalloc failed for the B assembly batches");
    }
    P->bq = bq; P->bgt = bgt; P->bt1 = bt1; P->bG = bG; P->bpsi = bpsi; P->bedge = bedge; P->Bd = Bd; P->bk = K;
    launch_fill(bG, (long)K * L * nG, 0.0, P->stream);  // the coarse-level rings stay zero
    P->Bh.assign((size_t)nf * nf, 0.0);
    int rc = assemble_b(P);
    if (rc != IBPM_OK) {
        release();
        return rc;
    }
    // inv(B) on the device: Gauss-Jordan with row pivoting on [B | I]
    double *M = nullptr, *pv = nullptr, *f = nullptr;
    int* iv = nullptr;
    e = hipMalloc(&M, sizeof(double) * 2 * (size_t)nf * nf);
    if (e == hipSuccess) e = hipMalloc(&pv, sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&f, sizeof(double) * nf);
    if (e == hipSuccess) e = hipMalloc(&iv, 2 * sizeof(int));
    std::vector<double> eye((size_t)nf * nf, 0.0);
    for (int i = 0; i < nf; i++) eye[(size_t)i + (size_t)nf * i] = 1.0;
    int hv[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(M, Bd, sizeof(double) * nf * nf, hipMemcpyDeviceToDevice, P->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(M + (size_t)nf * nf, eye.data(), sizeof(double) * nf * nf, hipMemcpyHostToDevice, P->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(iv, hv, 2 * sizeof(int), hipMemcpyHostToDevice, P->stream);
    if (e == hipSuccess) {
        launch_gauss_jordan(M, nf, iv, pv, iv + 1, f, P->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(eye.data(), M + (size_t)nf * nf, sizeof(double) * nf * nf,
                                            hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hv, iv, 2 * sizeof(int), hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
    for (void* p : {(void*)M, (void*)pv, (void*)f, (void*)iv}) if (p) (void)hipFree(p);
    release();
    if (e != hipSuccess) return fail(C, IBPM_E_HIP, std::string("inverting B: ") + hipGetErrorString(e));
    if (hv[1]) return fail(C, IBPM_E_SINGULAR, "B is singular");
    std::vector<double>& B = P->Bh;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < nf; i++)
        for (int k = 0; k < nf; k++) {
            double d = B[(size_t)i + (size_t)nf * k] - B[(size_t)k + (size_t)nf * i];
            num += d * d;
            den += B[(size_t)i + (size_t)nf * k] * B[(size_t)i + (size_t)nf * k];
        }
    P->b_asym = den > 0 ? std::sqrt(num / den) : 0.0;
    P->binv.swap(eye);
    B.clear();
    B.shrink_to_fit();
    return IBPM_OK;
}

static int upload_binv(ibpm_prob* P)
{
    const int nf = P->nf;
    std::vector<double> T((size_t)nf * nf);
    for (int i = 0; i < nf; i++)
        for (int k = 0; k < nf; k++) T[(size_t)i * nf + k] = P->binv[(size_t)i + (size_t)nf * k];
    HIPTRY(P->ctx, hipMemcpy(P->BinvT, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice));
    return IBPM_OK;
}

static int zero_state(ibpm_prob* P)
{
    hipStream_t s = P->stream;
    const long nG = P->g.nG, nq = P->g.nq;
    launch_fill(P->q, nq * P->mg, 0.0, s);
    P->q_stale = false;
    launch_fill(P->q0, nq * P->mg, 0.0, s);
    launch_fill(P->G, nG * P->mg, 0.0, s);
    launch_fill(P->Gs, nG * P->mg, 0.0, s);
    launch_fill(P->psi, nG * P->mg, 0.0, s);
    launch_fill(P->nl[0], nG * P->mg, 0.0, s);
    launch_fill(P->nl[1], nG * P->mg, 0.0, s);
    launch_fill(P->Fb, P->nf, 0.0, s);
    launch_fill(P->fbv, P->nf, 0.0, s);
    launch_fill(P->CD, P->nbodies, 0.0, s);
    launch_fill(P->CL, P->nbodies, 0.0, s);
    P->nl1 = P->nl2 = 0;
    P->q0_uniform = true;
    P->q0vx = P->q0vy = 0.0;
    return IBPM_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int ibpm_ctx_create(int device, ibpm_ctx** out)
{
    if (!out) return IBPM_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return IBPM_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return IBPM_E_HIP;
    kernels_init();
    void* sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) return IBPM_E_HIP;
    ibpm_ctx* c = new ibpm_ctx();
    c->device = device;
    c->scratch = static_cast<double*>(sp);
    *out = c;
    return IBPM_OK;
}

int ibpm_ctx_destroy(ibpm_ctx* ctx)
{
    delete ctx;
    return IBPM_OK;
}

const char* ibpm_last_error(const ibpm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ibpm_problem_create(ibpm_ctx* ctx, const ibpm_grid_desc* gd, const ibpm_body_desc* bd, const double* binv,
                        ibpm_prob** out)
{
    if (!ctx || !gd || !bd || !out) return IBPM_E_ARG;
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    HIPTRY(ctx, hipSetDevice(ctx->device));
    const int nx = gd->nx, ny = gd->ny, mg = gd->mg;
    if (nx < 4 || ny < 4 || mg < 1 || mg > MAXLEV)
        return fail(ctx, IBPM_E_ARG, "grid needs nx, ny >= 4 and 1 <= mg <= 16");
    if (mg > 1 && (nx % 4 != 0 || ny % 4 != 0 || nx < 8 || ny < 8))
        return fail(ctx, IBPM_E_UNSUPPORTED, "multi-level grids need nx, ny >= 8 and divisible by 4 (get_bc!/coarsify!)");
    if (bd->motion == IBPM_MOTION_FUNCTION && binv)
        return fail(ctx, IBPM_E_ARG, "MotionFunction problems have no Binv matrix (CG every step, laplacian.jl:123-141)");
    if (bd->motion == IBPM_MOTION_MOVING_GRID && bd->nbodies != 1)
        return fail(ctx, IBPM_E_ARG, "MovingGrid needs exactly one body (timestepping.jl:277)");
    ibpm_prob* P = new ibpm_prob();
    P->ctx = ctx;
    GridDev& g = P->g;
    g.nx = nx; g.ny = ny; g.n1 = nx - 1; g.n2 = ny - 1;
    g.nG = (long)(nx - 1) * (ny - 1);
    g.nu = (long)ny * (nx + 1); g.nv = (long)nx * (ny + 1); g.nq = g.nu + g.nv;
    P->xreuse = xreuse_default(nx, ny);
    g.LEFT = 0; g.RIGHT = ny + 1; g.BOT = 2 * (ny + 1); g.TOP = 2 * (ny + 1) + nx + 1;
    g.nbc = 2 * (nx + 1) + 2 * (ny + 1);
    g.h = gd->h;
    P->mg = mg; P->Re = gd->Re; P->dt = gd->dt; P->beta[0] = gd->beta[0]; P->beta[1] = gd->beta[1];
    P->offx = gd->offx; P->offy = gd->offy;
    P->motion = bd->motion;
    P->nbodies = bd->nbodies;
    P->boff.push_back(0);
    for (int b = 0; b < bd->nbodies; b++) {
        P->nb.push_back(bd->nb[b]);
        P->boff.push_back(P->boff.back() + bd->nb[b]);
    }
    P->nbt = P->boff.back();
    P->nf = 2 * P->nbt;
    const int nbt = P->nbt;
    P->xb.assign(bd->xb, bd->xb + 2 * (size_t)nbt);
    P->ds.assign(bd->ds, bd->ds + (size_t)nbt);
    P->x0 = P->xb;  // RigidBody.x0 = copy(xb) (sample-bodies.jl:32)
    P->ub.assign(2 * (size_t)nbt, 0.0);
    const bool mf = P->motion == IBPM_MOTION_FUNCTION;
    auto bail = [&](int code, const std::string& m) {
        ibpm_problem_destroy(P);
        return fail(ctx, code, m);
    };
    // ---- setup_reg (interface-coupling.jl:72-96)
    if (!setup_reg_host(P)) return bail(IBPM_E_ARG, "body point closer than 6 cells to the domain edge");
    // ---- coarsify! ranges (mg-utils.jl:77-80, zip truncates to the shorter range)
    {
        CoarsifyRange r{};
        int j0 = ny / 2 + ((-ny) / 2 + 2), j1 = ny / 2 + (ny / 2 - 2);
        int jc0 = ny / 2 + ((-ny) / 4 + 1), jc1 = ny / 2 + (ny / 4 - 1);
        int i0 = nx / 2 + ((-nx) / 2 + 2), i1 = nx / 2 + (nx / 2 - 2);
        int ic0 = nx / 2 + ((-nx) / 4 + 1), ic1 = nx / 2 + (nx / 4 - 1);
        int nj = j1 >= j0 ? (j1 - j0) / 2 + 1 : 0, njc = jc1 >= jc0 ? jc1 - jc0 + 1 : 0;
        int ni = i1 >= i0 ? (i1 - i0) / 2 + 1 : 0, nic = ic1 >= ic0 ? ic1 - ic0 + 1 : 0;
        r.i0 = i0; r.ic0 = ic0; r.ci = std::min(ni, nic);
        r.j0 = j0; r.jc0 = jc0; r.cj = std::min(nj, njc);
        P->crange.push_back(r);
    }
    // ---- DST plans and eigenvalue tables
    std::vector<double2> twx, wmx, twy, wmy;
    std::vector<double> snx, sny;
    if (!make_plan(nx, P->tx.plan, twx, snx, wmx) || !make_plan(ny, P->ty.plan, twy, sny, wmy))
        return bail(IBPM_E_UNSUPPORTED, "DST length unsupported (need nx, ny <= 1024, even with nx/2, ny/2 of the form 2^a 3^b 5^c for the FFT path)");
    std::vector<double> cxh(nx - 1), cyh(ny - 1);
    const long double fpi = (long double)(double)PI_L;
    for (int i = 1; i <= nx - 1; i++) cxh[i - 1] = std::cos((double)(fpi * (long double)i / (long double)nx));
    for (int j = 1; j <= ny - 1; j++) cyh[j - 1] = std::cos((double)(fpi * (long double)j / (long double)ny));
    // ---- device allocations
    hipError_t e = hipStreamCreate(&P->stream);
    if (e != hipSuccess) return bail(IBPM_E_HIP, "hipStreamCreate failed");
    const long nG = g.nG, nq = g.nq;
#define DA(ptr, n)                                                              \
    if ((e = dalloc(P, &(ptr), (size_t)(n))) != hipSuccess)                     \
        return bail(IBPM_E_HIP, std::string("hipMalloc failed for " #ptr));
    DA(P->q, nq * mg); DA(P->q0, nq * mg); DA(P->G, nG * mg); DA(P->Gs, nG * mg); DA(P->psi, nG * mg);
    DA(P->nl[0], nG * mg); DA(P->nl[1], nG * mg);
    DA(P->qs1, nq); DA(P->T1, nG * mg); DA(P->T2, nG * mg); DA(P->Gtmp, nG); DA(P->qreg, nq);
    DA(P->edge, 2 * (long)g.n2);
    DA(P->Fb, P->nf); DA(P->Ftmp, P->nf); DA(P->fbv, P->nf); DA(P->CD, P->nbodies); DA(P->CL, P->nbodies);
    DA(P->scal, 16); DA(P->slot, 1);
    DA(P->BinvT, mf ? 1 : (size_t)P->nf * P->nf); DA(P->w, P->weight.size()); DA(P->Fw, P->nf);
    DA(P->ds_d, nbt); DA(P->bidx, 2 * nbt); DA(P->boff_d, P->boff.size());
    // E' lists sized for any body position: at most one cell per window entry
    const size_t nent = (size_t)2 * nbt * 169;
    DA(P->sc_cell, nent); DA(P->sc_off, nent + 1); DA(P->sc_f, nent); DA(P->sc_w, nent);
    DA(P->cx, nx - 1); DA(P->cy, ny - 1);
    DA(P->tx.tw, twx.size()); DA(P->tx.sn, snx.size()); DA(P->tx.wm, wmx.size());
    DA(P->ty.tw, twy.size()); DA(P->ty.sn, sny.size()); DA(P->ty.wm, wmy.size());
#undef DA
    auto up = [&](void* d, const void* h_, size_t bytes) -> bool {
        return bytes == 0 || hipMemcpy(d, h_, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = upload_reg(P) == hipSuccess && hipStreamSynchronize(P->stream) == hipSuccess &&
              up(P->ds_d, P->ds.data(), sizeof(double) * P->ds.size()) &&
              up(P->boff_d, P->boff.data(), sizeof(int) * P->boff.size()) &&
              up(P->cx, cxh.data(), sizeof(double) * cxh.size()) && up(P->cy, cyh.data(), sizeof(double) * cyh.size()) &&
              up(P->tx.tw, twx.data(), sizeof(double2) * twx.size()) && up(P->tx.sn, snx.data(), sizeof(double) * snx.size()) &&
              up(P->tx.wm, wmx.data(), sizeof(double2) * wmx.size()) && up(P->ty.tw, twy.data(), sizeof(double2) * twy.size()) &&
              up(P->ty.sn, sny.data(), sizeof(double) * sny.size()) && up(P->ty.wm, wmy.data(), sizeof(double2) * wmy.size());
    if (!ok) return bail(IBPM_E_HIP, "upload of setup tables failed");
    P->tx.plan.tw = P->tx.tw; P->tx.plan.sn = P->tx.sn; P->tx.plan.wm = P->tx.wm;
    P->ty.plan.tw = P->ty.tw; P->ty.plan.sn = P->ty.sn; P->ty.plan.wm = P->ty.wm;
    // scratch fields whose untouched entries must stay zero
    launch_fill(P->qreg, nq, 0.0, P->stream);
    launch_fill(P->Gtmp, nG, 0.0, P->stream);
    launch_fill(P->edge, 2L * g.n2, 0.0, P->stream);
    zero_state(P);
    if (mf && P->nf > 0) {
        // pinned staging of the force solve's right-hand side and F̃b
        if (hipHostMalloc((void**)&P->pin, sizeof(double) * 2 * (size_t)P->nf, hipHostMallocDefault) != hipSuccess)
            return bail(IBPM_E_HIP, "hipHostMalloc failed for the CG staging");
        // batches of unit-force fields for the per-step B: as many as fit in 4 GiB
        const size_t per = sizeof(double) * ((size_t)nq + 2 * (size_t)nG + 2 * (size_t)mg * nG);
        // (at most 32768 fields: the batch is a grid dimension of launch_rot_batch)
        P->bk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)P->nf, ((size_t)4 << 30) / per, (size_t)32768}));
        const size_t K = (size_t)P->bk;
#define DB(ptr, n)                                                              \
        if ((e = dalloc(P, &(ptr), (size_t)(n))) != hipSuccess)                 \
            return bail(IBPM_E_HIP, std::string("hipMalloc failed for " #ptr));
        DB(P->bq, K * nq); DB(P->bgt, K * nG); DB(P->bt1, K * nG); DB(P->bG, K * mg * nG); DB(P->bpsi, K * mg * nG);
        DB(P->bedge, K * 2 * (size_t)g.n2);
        DB(P->Bd, (size_t)P->nf * P->nf);
#undef DB
        P->Bh.assign((size_t)P->nf * P->nf, 0.0);
        launch_fill(P->bG, (long)(K * mg * nG), 0.0, P->stream);  // the coarse-level rings stay zero
    }
    // ---- Binv (none for MotionFunction: CG every step)
    if (P->nf > 0 && !mf) {
        if (binv) {
            P->binv.assign(binv, binv + (size_t)P->nf * P->nf);
        } else {
            int rc = build_binv(P);
            if (rc != IBPM_OK) { std::string m = ctx->err; ibpm_problem_destroy(P); return fail(ctx, rc, m); }
        }
        int rc = upload_binv(P);
        if (rc != IBPM_OK) { std::string m = ctx->err; ibpm_problem_destroy(P); return fail(ctx, rc, m); }
    }
    zero_state(P);
    if (hipStreamSynchronize(P->stream) != hipSuccess) return bail(IBPM_E_HIP, "setup failed on the device");
    P->setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = P;
    return IBPM_OK;
}

int ibpm_problem_destroy(ibpm_prob* P)
{
    if (!P) return IBPM_OK;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: destroy the group first");
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    graphs_clear(P);
    P->hb.release();
    adj_free(P);
    for (void* p : P->allocs) (void)hipFree(p);
    if (P->pin) (void)hipHostFree(P->pin);
    for (hipEvent_t ev : P->pev) (void)hipEventDestroy(ev);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return IBPM_OK;
}

int ibpm_problem_linearize(ibpm_prob* P, const double* qB, const double* q0B, const double* GammaB)
{
    if (!P || !qB || !q0B || !GammaB) return IBPM_E_ARG;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: step or change it through the group");
    if (int rc = check_idle(P)) return rc;
    if (P->motion != IBPM_MOTION_STATIC)
        return fail(P->ctx, IBPM_E_UNSUPPORTED, "LinearizedIBProblem is defined for static bodies only");
    const GridDev& g = P->g;
    const int nx = g.nx, ny = g.ny, mg = P->mg;
    const long nq = g.nq, nG = g.nG, nu = g.nu;
    // QB[:, lev] = avg_flux(base_state, base_prob; lev) -- nonlinear.jl:17-38,
    // with the reference's order of the four-term sums; other entries zero
    std::vector<double> QB((size_t)(nq * mg), 0.0);
    for (int lev = 0; lev < mg; lev++) {
        const double* q = qB + (long)lev * nq;
        const double* q0 = q0B + (long)lev * nq;
        double* Q = QB.data() + (long)lev * nq;
        for (int j = 2; j <= ny; j++)
            for (int i = 1; i <= nx + 1; i++) {
                const long p = (i - 1) + (long)(nx + 1) * (j - 1), pm = p - (nx + 1);
                Q[p] = ((q[p] + q0[p]) + q[pm]) + q0[pm];
            }
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 2; i <= nx; i++) {
                const long p = nu + (i - 1) + (long)nx * (j - 1), pm = p - 1;
                Q[p] = ((q[p] + q0[p]) + q[pm]) + q0[pm];
            }
    }
    if (!P->QB) HIPTRY(P->ctx, dalloc(P, &P->QB, (size_t)(nq * mg)));
    if (!P->GB) HIPTRY(P->ctx, dalloc(P, &P->GB, (size_t)(nG * mg)));
    HIPTRY(P->ctx, hipMemcpyAsync(P->QB, QB.data(), sizeof(double) * QB.size(), hipMemcpyHostToDevice, P->stream));
    HIPTRY(P->ctx, hipMemcpyAsync(P->GB, GammaB, sizeof(double) * (size_t)(nG * mg), hipMemcpyHostToDevice, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    P->lin = true;
    graphs_clear(P);
    ibpm_step_scalars z{};
    return ibpm_base_flux(P, &z);
}

int ibpm_problem_get_info(const ibpm_prob* P, ibpm_problem_info* info)
{
    if (!P || !info) return IBPM_E_ARG;
    info->nx = P->g.nx; info->ny = P->g.ny; info->mg = P->mg; info->nbodies = P->nbodies; info->nb_total = P->nbt;
    info->nbc = P->g.nbc; info->nG = P->g.nG; info->nq = P->g.nq; info->h = P->g.h; info->offx = P->offx;
    info->offy = P->offy; info->b_asymmetry = P->b_asym; info->setup_seconds = P->setup_s;
    return IBPM_OK;
}

int ibpm_problem_get_binv(ibpm_prob* P, double* out, size_t n)
{
    if (P && P->motion == IBPM_MOTION_FUNCTION)
        return fail(P->ctx, IBPM_E_UNSUPPORTED, "MotionFunction problems have no Binv matrix (CG every step)");
    if (!P || !out || n != (size_t)P->nf * P->nf) return fail(P ? P->ctx : nullptr, IBPM_E_ARG, "binv size");
    std::copy(P->binv.begin(), P->binv.end(), out);
    return IBPM_OK;
}

int ibpm_problem_get_body_idx(ibpm_prob* P, int32_t* out, size_t n)
{
    if (!P || !out || n != 2 * (size_t)P->nbt) return fail(P ? P->ctx : nullptr, IBPM_E_ARG, "body_idx size");
    std::copy(P->body_idx.begin(), P->body_idx.end(), out);
    return IBPM_OK;
}

static double* field_ptr(ibpm_prob* P, int field, size_t* n, bool for_upload)
{
    const size_t nq = P->g.nq * P->mg, nG = P->g.nG * P->mg;
    switch (field) {
    case IBPM_FIELD_Q: *n = nq; return P->q;
    case IBPM_FIELD_Q0:
        *n = nq;
        if (for_upload) P->q0_uniform = false;
        return P->q0;
    case IBPM_FIELD_GAMMA: *n = nG; return P->G;
    case IBPM_FIELD_PSI: *n = nG; return P->psi;
    case IBPM_FIELD_NONLIN1:
        *n = nG;
        if (for_upload && P->nl1 == P->nl2) P->nl1 = 1 - P->nl2;
        return P->nl[P->nl1];
    case IBPM_FIELD_NONLIN2:
        *n = nG;
        if (for_upload && P->nl1 == P->nl2) P->nl2 = 1 - P->nl1;
        return P->nl[P->nl2];
    case IBPM_FIELD_FB_TILDE: *n = P->nf; return P->Fb;
    case IBPM_FIELD_FB: *n = P->nf; return P->fbv;
    case IBPM_FIELD_CD: *n = P->nbodies; return P->CD;
    case IBPM_FIELD_CL: *n = P->nbodies; return P->CL;
    default: return nullptr;
    }
}

int ibpm_state_upload(ibpm_prob* P, int field, const double* host, size_t n)
{
    if (!P || !host) return IBPM_E_ARG;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: step or change it through the group");
    size_t want = 0;
    double* d = field_ptr(P, field, &want, true);
    if (!d) return fail(P->ctx, IBPM_E_ARG, "unknown field");
    if (n != want) return fail(P->ctx, IBPM_E_ARG, "upload size mismatch");
    if (int rc = check_idle(P)) return rc;
    graphs_clear(P);
    ensure_q(P);  // the state's q before any of its fields changes
    HIPTRY(P->ctx, hipMemcpyAsync(d, host, sizeof(double) * n, hipMemcpyHostToDevice, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}

int ibpm_state_download(ibpm_prob* P, int field, double* host, size_t n)
{
    if (!P || !host) return IBPM_E_ARG;
    size_t want = 0;
    double* d = field_ptr(P, field, &want, false);
    if (!d) return fail(P->ctx, IBPM_E_ARG, "unknown field");
    if (n != want) return fail(P->ctx, IBPM_E_ARG, "download size mismatch");
    if (field == IBPM_FIELD_Q) ensure_q(P);
    HIPTRY(P->ctx, hipMemcpyAsync(host, d, sizeof(double) * n, hipMemcpyDeviceToHost, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}

int ibpm_base_flux(ibpm_prob* P, const ibpm_step_scalars* s)
{
    if (!P || !s) return IBPM_E_ARG;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: step or change it through the group");
    if (int rc = check_idle(P)) return rc;
    graphs_clear(P);  // captured nonlinear launches hold the base flux by value
    if (P->lin) {
        // static bodies of a linearized problem: q0 = 0 (timestepping.jl:238-250)
        launch_base_flux_inertial(P->q0, P->mg, 0.0, 0.0, P->g, P->stream);
        P->q0_uniform = true;
        P->q0vx = P->q0vy = 0.0;
        HIPTRY(P->ctx, hipGetLastError());
        return IBPM_OK;
    }
    if (P->motion == IBPM_MOTION_MOVING_GRID) {
        double Om = -s->theta_dot, al = -s->theta;
        double Ux0 = s->U * std::cos(al) - s->V * std::sin(al);
        double Uy0 = s->U * std::sin(al) + s->V * std::cos(al);
        Ux0 += s->Ux * std::cos(s->inclination) - s->Uy * std::sin(s->inclination);
        Uy0 += s->Ux * std::sin(s->inclination) + s->Uy * std::cos(s->inclination);
        BaseFluxArgs b{};
        b.q0 = P->q0; b.mg = P->mg; b.Ux0 = Ux0; b.Uy0 = Uy0; b.Om = Om; b.offx = P->offx; b.offy = P->offy; b.g = P->g;
        launch_base_flux_moving(b, P->stream);
        P->q0_uniform = false;
    } else {
        double a = s->inclination;
        double vx = s->Ux * std::cos(a) - s->Uy * std::sin(a);
        double vy = s->Ux * std::sin(a) + s->Uy * std::cos(a);
        launch_base_flux_inertial(P->q0, P->mg, vx, vy, P->g, P->stream);
        // the same values k_base_flux_inertial writes
        P->q0_uniform = true;
        P->q0vx = vx;
        P->q0vy = vy;
    }
    HIPTRY(P->ctx, hipGetLastError());
    return IBPM_OK;
}

int ibpm_state_init(ibpm_prob* P, const ibpm_step_scalars* s)
{
    if (!P || !s) return IBPM_E_ARG;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: step or change it through the group");
    if (int rc = check_idle(P)) return rc;
    zero_state(P);
    return ibpm_base_flux(P, s);
}

int ibpm_vort2flux(ibpm_prob* P)
{
    if (!P) return IBPM_E_ARG;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: step or change it through the group");
    if (int rc = check_idle(P)) return rc;
    vort2flux_dev(P, P->G, P->psi, P->q, CURL_ALL);
    HIPTRY(P->ctx, hipGetLastError());
    return IBPM_OK;
}

int ibpm_advance(ibpm_prob* P, double t, const ibpm_step_scalars* s)
{
    (void)t;
    if (!P || !s) return IBPM_E_ARG;
    int done = 0;
    return ibpm_advance_n(P, 1, &t, s, nullptr, nullptr, &done);
}

int ibpm_advance_async(ibpm_prob* P, int nsteps, const ibpm_step_scalars* s)
{
    if (!P || !s || nsteps < 0) return IBPM_E_ARG;
    if (P->group) return fail(P->ctx, IBPM_E_ARG, "the problem belongs to a group: step or change it through the group");
    const int nb = std::max(P->nbodies, 1);
    const size_t want = (size_t)P->pending + (size_t)nsteps;
    if (want > P->hist_steps) {
        // the history buffer grows only between waits (its halves hold CD and CL)
        if (P->pending > 0)
            return fail(P->ctx, IBPM_E_ARG, "history full: call ibpm_advance_wait before enqueuing more steps");
        // No free here: hipFree would wait for the whole device, i.e. for the
        // steps other problems have queued on their own streams (HistBuf keeps
        // the outgrown buffers until the problem is destroyed).
        graphs_clear(P);  // captured steps hold the old history buffer
        HIPTRY(P->ctx, P->hb.reserve(2 * want * (size_t)nb));
        P->hist = P->hb.d;
        P->hist_steps = P->hb.n / (2 * (size_t)nb);
    }
    if (nsteps > 0) launch_set_int(P->slot, P->pending, P->stream);
    if (P->motion == IBPM_MOTION_FUNCTION) {
        // each step solves on the host: run them now, keep what finished
        for (int k = 0; k < nsteps; k++) {
            int rc = step_mf(P, s[k], P->hist, P->hist + P->hist_steps * nb);
            if (rc != IBPM_OK) return rc;
            P->pending++;
        }
        return IBPM_OK;
    }
    for (int k = 0; k < nsteps; k++) step_dev(P, s[k], P->hist, P->hist + P->hist_steps * nb);
    P->pending += nsteps;
    HIPTRY(P->ctx, hipGetLastError());
    return IBPM_OK;
}

int ibpm_advance_wait(ibpm_prob* P, double* cd_hist, double* cl_hist, int* steps_done)
{
    if (!P) return IBPM_E_ARG;
    if (steps_done) *steps_done = 0;
    const int nb = std::max(P->nbodies, 1);
    const size_t n = (size_t)P->pending;
    // into the pinned mirror (a straight DMA), then out of it on the host
    if (n > 0) {
        HIPTRY(P->ctx, hipMemcpyAsync(P->hb.h, P->hist, sizeof(double) * n * nb, hipMemcpyDeviceToHost, P->stream));
        HIPTRY(P->ctx, hipMemcpyAsync(P->hb.h + P->hist_steps * nb, P->hist + P->hist_steps * nb,
                                      sizeof(double) * n * nb, hipMemcpyDeviceToHost, P->stream));
    }
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    std::vector<double> cd(n * nb), cl(n * nb);
    if (n > 0) {
        std::copy(P->hb.h, P->hb.h + n * nb, cd.begin());
        std::copy(P->hb.h + P->hist_steps * nb, P->hb.h + P->hist_steps * nb + n * nb, cl.begin());
    }
    collect_profile(P);
    P->pending = 0;
    if (P->async_err) {
        P->async_err = false;
        return fail(P->ctx, IBPM_E_HIP, "launching a captured step failed");
    }
    int done = (int)n;
    for (size_t k = 0; k < n && done == (int)n; k++)
        for (int b = 0; b < P->nbodies; b++)
            if (!std::isfinite(cl[k * nb + b])) { done = (int)k + 1; break; }
    if (cd_hist && P->nbodies > 0) std::copy(cd.begin(), cd.end(), cd_hist);
    if (cl_hist && P->nbodies > 0) std::copy(cl.begin(), cl.end(), cl_hist);
    if (steps_done) *steps_done = done;
    return done < (int)n ? fail(P->ctx, IBPM_E_NONFINITE, "non-finite CL") : IBPM_OK;
}

int ibpm_advance_n(ibpm_prob* P, int nsteps, const double* t, const ibpm_step_scalars* s, double* cd_hist,
                   double* cl_hist, int* steps_done)
{
    (void)t;
    if (!P || !s || nsteps < 0) return IBPM_E_ARG;
    if (steps_done) *steps_done = 0;
    if (P->pending > 0) return fail(P->ctx, IBPM_E_ARG, "steps enqueued by ibpm_advance_async are not waited for");
    const int rc = ibpm_advance_async(P, nsteps, s);
    if (rc != IBPM_OK) {
        // a step that failed part way through (a MotionFunction body leaving
        // the domain): collect the steps that finished, keep the failure
        const std::string msg = P->ctx->err;
        if (P->pending > 0) (void)ibpm_advance_wait(P, cd_hist, cl_hist, steps_done);
        P->pending = 0;
        return fail(P->ctx, rc, msg);
    }
    return ibpm_advance_wait(P, cd_hist, cl_hist, steps_done);
}

// ---- tangent-linear and adjoint runs ------------------------------------------
static int hist_reserve(ibpm_prob* P, int nsteps)
{
    const int nb = std::max(P->nbodies, 1);
    if ((size_t)nsteps > P->hist_steps) {
        graphs_clear(P);
        HIPTRY(P->ctx, P->hb.reserve(2 * (size_t)nsteps * nb));
        P->hist = P->hb.d;
        P->hist_steps = P->hb.n / (2 * (size_t)nb);
    }
    launch_set_int(P->slot, 0, P->stream);
    return IBPM_OK;
}

// J = Σ_k wcd_k CD_k + wcl_k CL_k of the steps just run (one body)
static int hist_objective(ibpm_prob* P, int nsteps, const double* wcd, const double* wcl, double* J)
{
    HIPTRY(P->ctx, hipMemcpyAsync(P->hb.h, P->hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost, P->stream));
    HIPTRY(P->ctx, hipMemcpyAsync(P->hb.h + P->hist_steps, P->hist + P->hist_steps, sizeof(double) * nsteps,
                                  hipMemcpyDeviceToHost, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    double j = 0.0;
    for (int k = 0; k < nsteps; k++) j += wcd[k] * P->hb.h[k] + wcl[k] * P->hb.h[P->hist_steps + k];
    *J = j;
    return IBPM_OK;
}

int ibpm_tangent_n(ibpm_prob* P, int nsteps, const ibpm_step_scalars* s, const double* ds, const double* dGamma0,
                   const double* wcd, const double* wcl, double* J, double* dJ)
{
    if (!P || !s || !ds || !wcd || !wcl || !J || !dJ || nsteps < 1) return IBPM_E_ARG;
    if (int rc = adj_check(P)) return rc;
    if (int rc = adj_setup(P, nsteps, 1)) return rc;
    if (int rc = hist_reserve(P, nsteps)) return rc;
    AdjWork* W = P->adj;
    const long n = P->g.nG * P->mg;
    snap_store(P, 0);
    aux_state(P, true);
    if (dGamma0) HIPTRY(P->ctx, hipMemcpyAsync(W->dG, dGamma0, sizeof(double) * n, hipMemcpyHostToDevice, P->stream));
    else launch_fill(W->dG, n, 0.0, P->stream);
    launch_fill(W->dN[0], n, 0.0, P->stream);
    int n0 = 0;
    for (int k = 0; k < nsteps; k++) {
        tlm_step(P, s[k], ds + 7L * k, k, n0);
        n0 = 1 - n0;
        step_body(P, s[k], P->hist, P->hist + P->hist_steps);
    }
    HIPTRY(P->ctx, hipGetLastError());
    std::vector<double> sums(2 * (size_t)nsteps);
    HIPTRY(P->ctx, hipMemcpyAsync(sums.data(), W->sums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, P->stream));
    if (int rc = hist_objective(P, nsteps, wcd, wcl, J)) return rc;
    const double cD = 2.0 * P->g.h / P->dt;
    double d = 0.0;
    for (int k = 0; k < nsteps; k++) d += cD * (wcd[k] * sums[2 * k] + wcl[k] * sums[2 * k + 1]);
    *dJ = d;
    snap_restore(P, 0);
    aux_state(P, false);
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}

int ibpm_adjoint_n(ibpm_prob* P, int nsteps, const ibpm_step_scalars* s, const double* wcd, const double* wcl,
                   int nsnap, double* J, double* sbar, double* Gamma0bar, int64_t* forward_steps)
{
    if (!P || !s || !wcd || !wcl || !J || !sbar || nsteps < 1 || nsnap < 1) return IBPM_E_ARG;
    if (int rc = adj_check(P)) return rc;
    if (int rc = adj_setup(P, nsteps, nsnap)) return rc;
    if (int rc = hist_reserve(P, nsteps)) return rc;
    AdjWork* W = P->adj;
    const long n = P->g.nG * P->mg;
    snap_store(P, 0);
    aux_state(P, true);
    for (int k = 0; k < nsteps; k++) step_body(P, s[k], P->hist, P->hist + P->hist_steps);
    HIPTRY(P->ctx, hipGetLastError());
    if (int rc = hist_objective(P, nsteps, wcd, wcl, J)) return rc;
    launch_fill(W->Gb, n, 0.0, P->stream);
    launch_fill(W->Nb1, n, 0.0, P->stream);
    AdjRun R{P, s, wcd, wcl};
    R.fwd_steps = nsteps;
    rev(R, 0, nsteps, 0, nsnap - 1);
    HIPTRY(P->ctx, hipGetLastError());
    std::vector<double> sums(3 * (size_t)nsteps);
    HIPTRY(P->ctx, hipMemcpyAsync(sums.data(), W->sums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, P->stream));
    if (Gamma0bar) HIPTRY(P->ctx, hipMemcpyAsync(Gamma0bar, W->Gb, sizeof(double) * n, hipMemcpyDeviceToHost, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    for (int k = 0; k < nsteps; k++) mg_flux_adjoint(s[k], sums.data() + 3 * (size_t)k, sbar + 7 * (size_t)k);
    if (forward_steps) *forward_steps = R.fwd_steps;
    snap_restore(P, 0);
    aux_state(P, false);
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}

int ibpm_get_forces(ibpm_prob* P, double* CD, double* CL, double* Fb)
{
    if (!P) return IBPM_E_ARG;
    if (CD && P->nbodies) HIPTRY(P->ctx, hipMemcpyAsync(CD, P->CD, sizeof(double) * P->nbodies, hipMemcpyDeviceToHost, P->stream));
    if (CL && P->nbodies) HIPTRY(P->ctx, hipMemcpyAsync(CL, P->CL, sizeof(double) * P->nbodies, hipMemcpyDeviceToHost, P->stream));
    if (Fb && P->nf) HIPTRY(P->ctx, hipMemcpyAsync(Fb, P->Fb, sizeof(double) * P->nf, hipMemcpyDeviceToHost, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}

int ibpm_get_body_state(ibpm_prob* P, double* xb, double* ub, size_t n)
{
    if (!P || n != 2 * (size_t)P->nbt) return fail(P ? P->ctx : nullptr, IBPM_E_ARG, "body state size");
    if (xb) std::copy(P->xb.begin(), P->xb.end(), xb);
    if (ub) std::copy(P->ub.begin(), P->ub.end(), ub);
    return IBPM_OK;
}

int ibpm_cg_stats(ibpm_prob* P, int32_t* iterations, double* residual, double* tolerance)
{
    if (!P) return IBPM_E_ARG;
    if (iterations) *iterations = P->cg_iters;
    if (residual) *residual = P->cg_res;
    if (tolerance) *tolerance = P->cg_tol;
    return IBPM_OK;
}

int ibpm_op_update_bodies(ibpm_prob* P, const double* kin)
{
    if (!P || !kin) return IBPM_E_ARG;
    if (P->motion != IBPM_MOTION_FUNCTION)
        return fail(P->ctx, IBPM_E_ARG, "update_bodies! moves MotionFunction bodies only");
    if (int rc = check_idle(P)) return rc;
    graphs_clear(P);
    int rc = update_bodies_dev(P, kin);
    if (rc != IBPM_OK) return rc;
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}


// ---------------------------------------------------------------------------
// Groups: same-shape static problems stepped together, every pass of the step
// one launch over all instances (fields laid out instance-major: level l of
// instance i at (i L + l) nG, resp. nq)
// ---------------------------------------------------------------------------
}  // extern "C"

struct ibpm_group {
    ibpm_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    std::vector<ibpm_prob*> probs;
    int n = 0, L = 1;
    GridDev g{};
    DevTables tx, ty;  // the first member's plans (identical for all members)
    double *cx = nullptr, *cy = nullptr;
    std::vector<CoarsifyRange> crange;
    double *G = nullptr, *Gs = nullptr, *psi = nullptr, *nl[2] = {nullptr, nullptr}, *q = nullptr, *q0 = nullptr;
    double *qs1 = nullptr, *T1 = nullptr, *T2 = nullptr, *Gt = nullptr, *qreg = nullptr;
    double* edge = nullptr;  // per instance: the vort2flux! chain's get_bc! columns (see ibpm_prob::edge)
    int nl1 = 0, nl2 = 1;
    bool q0_uniform = true;
    // bodies: all instances' points and F̃b entries concatenated
    int NBT = 0, NF = 0;
    double *Fb = nullptr, *Ftmp = nullptr, *BinvT = nullptr, *w = nullptr;
    int *bidx = nullptr, *pinst = nullptr, *plocal = nullptr, *inbt = nullptr, *ifoff = nullptr, *inf = nullptr;
    int *rinst = nullptr, *rloc = nullptr, *slots = nullptr;
    long* aoff = nullptr;
    InstPar* ipar = nullptr;
    ForceArgs* fargs = nullptr;
    std::vector<ForceArgs> fargs_h;
    int *sc_cell = nullptr, *sc_off = nullptr, *sc_f = nullptr;
    double* sc_w = nullptr;
    int sc_ncell = 0;
    int rb_i0 = 2, rb_i1 = 1, rb_j0 = 2, rb_j1 = 1;
    std::vector<int> nbh, hoff;  // history: instance i's CD block at hist + hoff[i], CL block after it
    HistBuf hb;              // history storage (device and pinned host mirror)
    double* hist = nullptr;  // = hb.d
    size_t hist_steps = 0;
    int pending = 0;
    int ytile = 4, ytile_single = 4, xwaves = 4, pipe = 1, ysplit = 1, xreuse = 0;
    // HIP graphs of one group step, one per buffer role (as ibpm_prob's);
    // every launch argument is fixed by the roles (the force arguments and
    // history pointers are read from device memory when the kernels run)
    bool graphs = true;
    // q_stale: the group's q is not formed (the steps leave it to ψ, as the
    // members' own FP64 steps do); its first step after creation reads q
    bool q_stale = false;
    // level-1 lines E reads of every member (their union; -1: not all interior)
    int tlo = -1, thi = -1;
    struct StepGraph {
        double* G;
        int out, nl2;
        bool qs;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> gcache;
    std::vector<void*> allocs;
    // the members' own buffers, restored when the group is destroyed
    struct Own {
        double *G, *Gs, *psi, *nl0, *nl1b, *q, *q0, *Fb;
        int nl1, nl2;
    };
    std::vector<Own> own;
};

namespace {

template <typename T>
hipError_t galloc(ibpm_group* Gp, T** p, size_t n)
{
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (n > 0 ? n : 1));
    if (e == hipSuccess) {
        Gp->allocs.push_back(v);
        *p = static_cast<T*>(v);
    }
    return e;
}

XArgs gxargs(const ibpm_group* Gp)
{
    XArgs a{};
    a.lstride = Gp->g.nG;
    a.mg = Gp->L;
    a.pl = Gp->tx.plan;
    a.g = Gp->g;
    a.zlo = 1;
    a.zhi = Gp->g.n2;
    a.waves = Gp->xwaves;
    a.ipar = Gp->ipar;
    return a;
}

YArgs gyargs(const ibpm_group* Gp, int op, double* buf)
{
    YArgs a{};
    a.buf = buf;
    a.op = op;
    a.cx = Gp->cx;
    a.cy = Gp->cy;
    a.scale = 1.0 / (double)(4L * Gp->g.nx * Gp->g.ny);
    a.sinx = Gp->tx.sn;
    a.bcfac = 1.0;
    a.ipar = Gp->ipar;
    a.pl = Gp->ty.plan;
    a.g = Gp->g;
    a.pipe = Gp->pipe;
    a.split = Gp->ysplit;
    return a;
}

// vort2flux! (curl.jl:133-165) of every instance: the level chain runs over
// all instances at once (one launch per pass and level)
// yhat, ysub0: as vort2flux_dev's (every instance's Ainv y-pass output, and
// the projection's y-pass output), so that a group's chains are an
// instance's bit for bit
// qout null: no curl! (the step's readers form the fluxes from ψ).  trial:
// as vort2flux_dev's CURL_TRIAL, only the ψ lines E and the chain read are
// transformed and stored (the members' level-1 lines tlo..thi, their union)
void group_vort2flux(ibpm_group* Gp, double* Gb, double* psib, double* qout, bool level1_only,
                     const double* yhat = nullptr, const double* ysub0 = nullptr, bool trial = false)
{
    hipStream_t s = Gp->stream;
    const long nG = Gp->g.nG, ist = (long)Gp->L * nG;
    const int L = Gp->L, n = Gp->n;
    for (int lev = 2; lev <= L; lev++) {
        CoarsifyArgs c{};
        const CoarsifyRange& r = Gp->crange[0];
        c.Gc = Gb + (long)(lev - 1) * nG;
        c.G = Gb + (long)(lev - 2) * nG;
        c.i0 = r.i0; c.ic0 = r.ic0; c.ci = r.ci; c.j0 = r.j0; c.jc0 = r.jc0; c.cj = r.cj;
        c.nb = n; c.bstride = ist; c.g = Gp->g; c.ipar = Gp->ipar; c.clev = lev;
        launch_coarsify(c, s);
    }
    // forward x-transform of every level of every instance in one launch,
    // then the chain (boundary terms in the y-pass, as in vort2flux_dev)
    XArgs a = gxargs(Gp);
    a.src = Gb; a.dst = Gp->T1; a.nlev = n * L; a.lev0 = 0; a.levmod = L;
    // with yhat only the centre rows coarsify! writes on levels 2..L (level
    // 1 keeps the projection's output that ysub0 points at)
    const int clo = Gp->crange[0].jc0 - 1, chi = Gp->crange[0].jc0 + Gp->crange[0].cj - 2;
    if (yhat) {
        a.lsel = 1; a.llo = clo; a.lhi = chi; a.minlev = 1;
    }
    if (!yhat || (L > 1 && chi >= clo)) launch_xpass(X_PLAIN, a, s);
    const bool rtrial = trial && !qout && Gp->tlo >= 0;
    for (int l = L - 1; l >= 0; l--) {
        YArgs y = gyargs(Gp, Y_DINV, Gp->T1 + (long)l * nG);
        y.lstride = ist; y.nlev = n; y.lev0 = l; y.fixlev = 1;
        if (rtrial) {  // the rows the trial's x-pass of this level reads (below)
            y.wsel = 1;
            y.wlo = l > 0 ? Gp->g.ny / 4 - 1 : Gp->tlo;
            y.whi = l > 0 ? 3 * Gp->g.ny / 4 - 1 : Gp->thi;
        }
        if (yhat) {
            y.ysrc = yhat + (long)l * nG;
            y.afac = 2.0 * (double)Gp->g.nx;
            y.clo = l > 0 ? clo : 1;
            y.chi = l > 0 ? chi : 0;
            if (l == 0) y.ysub = ysub0;
        }
        if (l < L - 1) {  // (an instance whose coarsest level is l has none: the kernel checks)
            y.bcr = psib + (long)(l + 1) * nG;
            y.bcstride = ist;
            y.edge = Gp->edge;
            y.edstride = 2L * Gp->g.n2;
        }
        launch_ypass(y, Gp->ytile_single, s);
        XArgs c = gxargs(Gp);
        c.src = Gp->T1 + (long)l * nG; c.dst = psib + (long)l * nG; c.lstride = ist; c.nlev = n; c.lev0 = l;
        c.fixlev = 1;
        if (l > 0) {
            c.edge = Gp->edge;
            c.edstride = 2L * Gp->g.n2;
        }
        if (rtrial) {
            c.lsel = 1;
            if (l > 0) {
                c.llo = c.keep0 = Gp->g.ny / 4 - 1;
                c.lhi = c.keep1 = 3 * Gp->g.ny / 4 - 1;
                c.keep = 1;
            } else {
                c.llo = Gp->tlo;
                c.lhi = Gp->thi;
            }
        }
        launch_xpass(X_PLAIN, c, s);
    }
    if (!qout) return;
    CurlArgs ca{};
    ca.psi = psib; ca.q = qout; ca.qstride = Gp->g.nq; ca.mg = L; ca.g = Gp->g; ca.lev0 = 0; ca.ipar = Gp->ipar;
    if (level1_only) {
        ca.nlev = n; ca.colbatch = 1; ca.psl = nG; ca.psb = ist;
    } else {
        ca.nlev = n * L; ca.levmod = L;
    }
    launch_curl(ca, s);
}

// the host side of a group step: the members follow the group's buffer roles,
// and q is left to ψ (formed where something reads it, ensure_q)
static void group_roles(ibpm_group* Gp)
{
    const long ist = (long)Gp->L * Gp->g.nG;
    Gp->q_stale = true;
    for (int i = 0; i < Gp->n; i++) {
        ibpm_prob* P = Gp->probs[i];
        P->G = Gp->G + (long)i * ist;
        P->Gs = Gp->Gs + (long)i * ist;
        P->nl1 = Gp->nl1;
        P->nl2 = Gp->nl2;
        P->q_stale = true;
    }
}

// advance!(state, prob, t) of every instance (timestepping.jl:35-67), static bodies
void group_step(ibpm_group* Gp)
{
    hipStream_t s = Gp->stream;
    const GridDev& g = Gp->g;
    const long nG = g.nG, ist = (long)Gp->L * nG;
    const int L = Gp->L, n = Gp->n, NE = n * L;
    // ---- get_trial_state!: all levels of all instances per launch
    const int out = (Gp->nl1 != Gp->nl2) ? Gp->nl1 : 1 - Gp->nl2;
    NLArgs na{};
    na.q = Gp->q; na.q0 = Gp->q0; na.G = Gp->G; na.out = Gp->nl[out];
    na.psi = Gp->q_stale ? Gp->psi : nullptr;  // the fluxes from ψ once the steps leave q unformed
    na.nlev = NE; na.lev0 = 0; na.mg = L; na.g = g; na.levmod = L; na.ipar = Gp->ipar;
    na.q0_uniform = Gp->q0_uniform ? 1 : 0;
    launch_nonlinear(na, s);
    // rhs = A Γ + ... (A by its 5-point stencil) and Ainv's first x-transform in one pass
    XArgs t = gxargs(Gp);
    t.src = Gp->G; t.dst = Gp->T2; t.nlev = NE; t.lev0 = 0;
    t.n1 = Gp->nl[out]; t.n0 = Gp->nl[Gp->nl2]; t.gam = Gp->G; t.levmod = L; t.ipar = Gp->ipar;
    launch_xpass(X_TRIAL, t, s);
    YArgs yi = gyargs(Gp, Y_AINV, Gp->T2);
    yi.lstride = nG; yi.nlev = NE; yi.lev0 = 0; yi.levmod = L;
    launch_ypass(yi, Gp->ytile, s);
    XArgs c = gxargs(Gp);
    c.src = Gp->T2; c.dst = Gp->Gs; c.nlev = NE; c.lev0 = 0; c.levmod = L;
    launch_xpass(X_PLAIN, c, s);
    Gp->nl1 = Gp->nl2 = out;
    group_vort2flux(Gp, Gp->Gs, Gp->psi, nullptr, true, Gp->xreuse ? Gp->T2 : nullptr, nullptr, true);
    // ---- boundary_forces! and update_stress! of every instance; E forms the
    // trial flux qs[:,1] from the trial ψ (curl!'s values)
    GatherArgs ga{};
    ga.q = nullptr; ga.q0 = Gp->q0; ga.w = Gp->w; ga.bidx = Gp->bidx; ga.out = Gp->Ftmp; ga.nbt = Gp->NBT; ga.g = g;
    ga.psi = Gp->psi; ga.psic = L > 1 ? Gp->psi + nG : nullptr; ga.psistride = ist; ga.ipar = Gp->ipar;
    ga.qstride = g.nq; ga.q0stride = (long)L * g.nq;
    ga.pinst = Gp->pinst; ga.plocal = Gp->plocal; ga.inbt = Gp->inbt; ga.ifoff = Gp->ifoff;
    launch_gather(ga, s);
    launch_gemv_forces_group(Gp->BinvT, Gp->aoff, Gp->Ftmp, Gp->Fb, Gp->rinst, Gp->rloc, Gp->ifoff, Gp->inf, Gp->NF,
                             Gp->fargs, n, s);
    // ---- project_circ!: Γs[:, 1] -= Ainv rot(E' F̃b) of every instance
    ScatterArgs sa{};
    sa.cell = Gp->sc_cell; sa.off = Gp->sc_off; sa.ent_f = Gp->sc_f; sa.ent_w = Gp->sc_w;
    sa.fb = Gp->Fb; sa.q = Gp->qreg; sa.ncell = Gp->sc_ncell;
    launch_scatter(sa, s);
    launch_rot_batch(Gp->qreg, g.nq, Gp->Gt, ist, n, Gp->rb_i0, Gp->rb_i1, Gp->rb_j0, Gp->rb_j1, g, s);
    XArgs p1 = gxargs(Gp);
    p1.src = Gp->Gt; p1.dst = Gp->T1; p1.lstride = ist; p1.nlev = n; p1.lev0 = 0; p1.fixlev = 1;
    p1.zlo = Gp->rb_j0 - 1; p1.zhi = std::min(Gp->rb_j1 - 1, g.n2);
    launch_xpass(X_PLAIN, p1, s);
    YArgs yp = gyargs(Gp, Y_AINV, Gp->T1);
    yp.lstride = ist; yp.nlev = n; yp.lev0 = 0; yp.fixlev = 1;
    launch_ypass(yp, Gp->ytile_single, s);
    XArgs p2 = gxargs(Gp);
    p2.src = Gp->T1; p2.dst = Gp->Gs; p2.lstride = ist; p2.nlev = n; p2.lev0 = 0; p2.fixlev = 1;
    launch_xpass(X_SUB, p2, s);
    std::swap(Gp->G, Gp->Gs);
    group_vort2flux(Gp, Gp->G, Gp->psi, nullptr, false, Gp->xreuse ? Gp->T2 : nullptr, Gp->xreuse ? Gp->T1 : nullptr);
    group_roles(Gp);
}

// history blocks of every instance for `steps` steps, and their pointers in
// the per-instance force arguments
int group_hist(ibpm_group* Gp, size_t steps)
{
    if (steps <= Gp->hist_steps) return IBPM_OK;
    const size_t cap = std::max({steps, 2 * Gp->hist_steps, (size_t)64});
    size_t tot = 0;
    for (int i = 0; i < Gp->n; i++) tot += 2 * cap * (size_t)Gp->nbh[i];
    HIPTRY(Gp->ctx, Gp->hb.reserve(tot));
    Gp->hist = Gp->hb.d;
    Gp->hist_steps = cap;
    tot = 0;
    for (int i = 0; i < Gp->n; i++) {
        Gp->hoff[i] = (int)tot;
        tot += 2 * cap * (size_t)Gp->nbh[i];
    }
    for (int i = 0; i < Gp->n; i++) {
        Gp->fargs_h[i].cd_hist = Gp->hist + Gp->hoff[i];
        Gp->fargs_h[i].cl_hist = Gp->hist + Gp->hoff[i] + cap * (size_t)Gp->nbh[i];
    }
    HIPTRY(Gp->ctx, hipMemcpyAsync(Gp->fargs, Gp->fargs_h.data(), sizeof(ForceArgs) * Gp->n, hipMemcpyHostToDevice,
                                   Gp->stream));
    // the staging copy must complete before fargs_h changes again
    HIPTRY(Gp->ctx, hipStreamSynchronize(Gp->stream));
    return IBPM_OK;
}

void group_graphs_clear(ibpm_group* Gp)
{
    for (auto& e : Gp->gcache) (void)hipGraphExecDestroy(e.exec);
    Gp->gcache.clear();
}

// group_step replayed from a graph captured for the current buffer roles
// (captured, instantiated and launched on first use of a role)
void group_step_dev(ibpm_group* Gp)
{
    if (!Gp->graphs) {
        group_step(Gp);
        return;
    }
    const int out = (Gp->nl1 != Gp->nl2) ? Gp->nl1 : 1 - Gp->nl2;
    for (auto& e : Gp->gcache) {
        if (e.G == Gp->G && e.out == out && e.nl2 == Gp->nl2 && e.qs == Gp->q_stale) {
            (void)hipGraphLaunch(e.exec, Gp->stream);
            // the host side of group_step: roles of the buffers
            Gp->nl1 = Gp->nl2 = out;
            std::swap(Gp->G, Gp->Gs);
            group_roles(Gp);
            return;
        }
    }
    ibpm_group::StepGraph e{Gp->G, out, Gp->nl2, Gp->q_stale, nullptr};
    double *G0 = Gp->G, *Gs0 = Gp->Gs;
    const int nl1 = Gp->nl1, nl2 = Gp->nl2;
    const bool qs0 = Gp->q_stale;
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(Gp->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        group_step(Gp);
        ok = hipStreamEndCapture(Gp->stream, &graph) == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (ok) ok = hipGraphLaunch(e.exec, Gp->stream) == hipSuccess;
    if (ok) {
        Gp->gcache.push_back(e);
        return;
    }
    // capture unavailable: undo the host side of the captured step, run it directly
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    (void)hipGetLastError();
    Gp->graphs = false;
    Gp->G = G0; Gp->Gs = Gs0; Gp->nl1 = nl1; Gp->nl2 = nl2; Gp->q_stale = qs0;
    for (int i = 0; i < Gp->n; i++) Gp->probs[i]->q_stale = qs0;
    group_step(Gp);
}

// Several groups stepped by one graph: one step of every group captured with
// a branch per group (each group's launches on its own stream, forked from and
// joined into the first group's stream), kept per combination of the groups'
// buffer roles.  The registry lets ibpm_group_destroy drop the graphs that
// name a destroyed group.
struct MultiGraph {
    std::vector<ibpm_group*> groups;
    std::vector<double*> G;
    std::vector<int> out, nl2, qs;
    hipGraphExec_t exec;
};
std::vector<MultiGraph> g_multi;
std::vector<hipEvent_t> g_multi_ev;

void multi_drop(ibpm_group* Gp)
{
    for (size_t k = 0; k < g_multi.size();) {
        auto& m = g_multi[k];
        if (std::find(m.groups.begin(), m.groups.end(), Gp) != m.groups.end()) {
            (void)hipGraphExecDestroy(m.exec);
            g_multi.erase(g_multi.begin() + k);
        } else {
            k++;
        }
    }
}

int multi_events(size_t n)
{
    while (g_multi_ev.size() < n) {
        hipEvent_t ev;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return IBPM_E_HIP;
        g_multi_ev.push_back(ev);
    }
    return IBPM_OK;
}

}  // namespace

extern "C" {

int ibpm_group_create(ibpm_ctx* ctx, int n, ibpm_prob* const* probs, ibpm_group** out)
{
    if (!ctx || !probs || !out || n < 1) return IBPM_E_ARG;
    *out = nullptr;
    const ibpm_prob* P0 = probs[0];
    for (int i = 0; i < n; i++) {
        const ibpm_prob* P = probs[i];
        if (!P || P->ctx != ctx) return fail(ctx, IBPM_E_ARG, "group members must be problems of this context");
        if (P->group) return fail(ctx, IBPM_E_ARG, "a problem can belong to one group only");
        if (P->motion != IBPM_MOTION_STATIC || P->lin || P->nbodies < 1)
            return fail(ctx, IBPM_E_UNSUPPORTED, "groups take problems with static bodies (not linearized)");
        if (P->g.nx != P0->g.nx || P->g.ny != P0->g.ny || P->g.h != P0->g.h || P->offx != P0->offx ||
            P->offy != P0->offy)
            return fail(ctx, IBPM_E_ARG, "group members must share the finest grid (nx, ny, h, offset)");
        if (P->pending > 0) return fail(ctx, IBPM_E_ARG, "wait for a member's queued steps before grouping it");
        for (int k = 0; k < i; k++)
            if (probs[k] == P) return fail(ctx, IBPM_E_ARG, "a problem appears twice in the group");
    }
    ibpm_group* Gp = new ibpm_group();
    Gp->ctx = ctx;
    Gp->n = n;
    Gp->L = 0;
    for (int i = 0; i < n; i++) Gp->L = std::max(Gp->L, probs[i]->mg);
    Gp->g = P0->g;
    Gp->tx = P0->tx; Gp->ty = P0->ty; Gp->cx = P0->cx; Gp->cy = P0->cy; Gp->crange = P0->crange;
    Gp->ytile = P0->ytile; Gp->ytile_single = P0->ytile_single; Gp->xwaves = P0->xwaves; Gp->pipe = P0->pipe; Gp->ysplit = P0->ysplit;
    Gp->xreuse = P0->xreuse;
    Gp->tlo = P0->tlo;
    Gp->thi = P0->thi;
    for (int i = 1; i < n && Gp->tlo >= 0; i++) {
        if (probs[i]->tlo < 0) Gp->tlo = Gp->thi = -1;
        else {
            Gp->tlo = std::min(Gp->tlo, probs[i]->tlo);
            Gp->thi = std::max(Gp->thi, probs[i]->thi);
        }
    }
    Gp->probs.assign(probs, probs + n);
    auto bail = [&](int code, const std::string& m) {
        for (void* p : Gp->allocs) (void)hipFree(p);
        if (Gp->stream) (void)hipStreamDestroy(Gp->stream);
        delete Gp;
        return fail(ctx, code, m);
    };
    for (int i = 0; i < n; i++) {
        ensure_q(probs[i]);  // the group's steps read q
        if (hipStreamSynchronize(probs[i]->stream) != hipSuccess) return bail(IBPM_E_HIP, "draining a member failed");
    }
    if (hipStreamCreate(&Gp->stream) != hipSuccess) return bail(IBPM_E_HIP, "hipStreamCreate failed");
    const long nG = Gp->g.nG, nq = Gp->g.nq, L = Gp->L;
    const size_t fG = (size_t)n * L * nG, fq = (size_t)n * L * nq;
    // instance maps of the bodies
    std::vector<int> pinst, plocal, inbt(n), ifoff(n), inf(n), rinst, rloc;
    std::vector<long> aoff(n);
    long atot = 0;
    for (int i = 0; i < n; i++) {
        const ibpm_prob* P = probs[i];
        inbt[i] = P->nbt;
        ifoff[i] = Gp->NF;
        inf[i] = P->nf;
        aoff[i] = atot;
        atot += (long)P->nf * P->nf;
        for (int k = 0; k < P->nbt; k++) { pinst.push_back(i); plocal.push_back(k); }
        for (int r = 0; r < P->nf; r++) { rinst.push_back(i); rloc.push_back(r); }
        Gp->NBT += P->nbt;
        Gp->NF += P->nf;
        Gp->nbh.push_back(std::max(P->nbodies, 1));
    }
    Gp->hoff.assign(n, 0);
    // E' of all instances: one segmented-gather list over the concatenated
    // fields (instances in order, each in its own cell order)
    std::vector<int> cell, off, ef;
    std::vector<double> ew;
    int ib0 = Gp->g.nx + 1, ib1 = 0, jb0 = Gp->g.ny + 2, jb1 = 0;
    for (int i = 0; i < n; i++) {
        const ibpm_prob* P = probs[i];
        const RegLists& R = P->reg;
        for (size_t c = 0; c < R.cell.size(); c++) {
            cell.push_back((int)(R.cell[c] + (long)i * nq));
            off.push_back((int)ef.size() + (R.off[c] - R.off[0]));
        }
        for (size_t e = 0; e < R.ef.size(); e++) {
            ef.push_back(R.ef[e] + ifoff[i]);
            ew.push_back(R.ew[e]);
        }
        ib0 = std::min(ib0, P->rb_i0); ib1 = std::max(ib1, P->rb_i1);
        jb0 = std::min(jb0, P->rb_j0); jb1 = std::max(jb1, P->rb_j1);
    }
    off.push_back((int)ef.size());
    Gp->sc_ncell = (int)cell.size();
    Gp->rb_i0 = ib0; Gp->rb_i1 = ib1; Gp->rb_j0 = jb0; Gp->rb_j1 = jb1;
    if ((size_t)n * nq > (size_t)INT32_MAX) return bail(IBPM_E_ARG, "group too large for 32-bit cell indices");
    hipError_t e;
#define GA(ptr, cnt)                                                             \
    if ((e = galloc(Gp, &(ptr), (size_t)(cnt))) != hipSuccess)                   \
        return bail(IBPM_E_HIP, std::string("hipMalloc failed for group " #ptr));
    GA(Gp->G, fG); GA(Gp->Gs, fG); GA(Gp->psi, fG); GA(Gp->nl[0], fG); GA(Gp->nl[1], fG);
    GA(Gp->q, fq); GA(Gp->q0, fq); GA(Gp->qs1, (size_t)n * nq); GA(Gp->T1, fG); GA(Gp->T2, fG); GA(Gp->Gt, fG);
    GA(Gp->qreg, (size_t)n * nq); GA(Gp->edge, (size_t)n * 2 * Gp->g.n2);
    GA(Gp->Fb, Gp->NF); GA(Gp->Ftmp, Gp->NF); GA(Gp->BinvT, atot); GA(Gp->w, (size_t)2 * Gp->NBT * 169);
    GA(Gp->bidx, 2 * (size_t)Gp->NBT); GA(Gp->pinst, Gp->NBT); GA(Gp->plocal, Gp->NBT); GA(Gp->inbt, n);
    GA(Gp->ifoff, n); GA(Gp->inf, n); GA(Gp->rinst, Gp->NF); GA(Gp->rloc, Gp->NF); GA(Gp->slots, n);
    GA(Gp->aoff, n); GA(Gp->ipar, n); GA(Gp->fargs, n);
    GA(Gp->sc_cell, cell.size()); GA(Gp->sc_off, off.size()); GA(Gp->sc_f, ef.size()); GA(Gp->sc_w, ew.size());
#undef GA
    hipStream_t s = Gp->stream;
    // levels a member does not have stay zero (they are never read)
    for (double* f : {Gp->G, Gp->Gs, Gp->psi, Gp->nl[0], Gp->nl[1]}) launch_fill(f, (long)fG, 0.0, s);
    for (double* f : {Gp->q, Gp->q0}) launch_fill(f, (long)fq, 0.0, s);
    auto h2d = [&](void* d, const void* h, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
    };
    auto d2d = [&](void* d, const void* src, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, s) == hipSuccess;
    };
    std::vector<InstPar> ip(n);
    Gp->fargs_h.resize(n);
    bool ok = true;
    for (int i = 0; i < n && ok; i++) {
        ibpm_prob* P = probs[i];
        const long io = (long)i * L;
        ip[i] = InstPar{P->beta[0] * P->dt, P->beta[1] * P->dt, P->dt, P->Re, P->q0vx, P->q0vy, P->mg, 0};
        const long Li = P->mg;
        Gp->q0_uniform = Gp->q0_uniform && P->q0_uniform;
        // the member's state into its slices; nonlin[1] and nonlin[2] as roles 0 and 1
        ok = d2d(Gp->G + io * nG, P->G, sizeof(double) * Li * nG) && d2d(Gp->psi + io * nG, P->psi, sizeof(double) * Li * nG) &&
             d2d(Gp->nl[0] + io * nG, P->nl[P->nl1], sizeof(double) * Li * nG) &&
             d2d(Gp->nl[1] + io * nG, P->nl[P->nl2], sizeof(double) * Li * nG) &&
             d2d(Gp->q + io * nq, P->q, sizeof(double) * Li * nq) && d2d(Gp->q0 + io * nq, P->q0, sizeof(double) * Li * nq) &&
             d2d(Gp->Fb + ifoff[i], P->Fb, sizeof(double) * P->nf) &&
             d2d(Gp->BinvT + aoff[i], P->BinvT, sizeof(double) * P->nf * P->nf);
        // weights and indices of instance i's points at its global point offset
        const long po = ifoff[i] / 2;  // points before instance i
        ok = ok && d2d(Gp->w + po * 2 * 169, P->w, sizeof(double) * 2 * 169 * P->nbt) &&
             d2d(Gp->bidx + po, P->bidx, sizeof(int) * P->nbt) &&
             d2d(Gp->bidx + Gp->NBT + po, P->bidx + P->nbt, sizeof(int) * P->nbt);
        ForceArgs& fa = Gp->fargs_h[i];
        fa = ForceArgs{};
        fa.Fb = Gp->Fb + ifoff[i]; fa.ds = P->ds_d; fa.boff = P->boff_d; fa.fb = P->fbv; fa.CD = P->CD; fa.CL = P->CL;
        fa.slot = Gp->slots + i; fa.nbh = Gp->nbh[i]; fa.nbodies = P->nbodies; fa.nbt = P->nbt; fa.hdt = Gp->g.h / P->dt;
    }
    ok = ok && h2d(Gp->pinst, pinst.data(), sizeof(int) * pinst.size()) &&
         h2d(Gp->plocal, plocal.data(), sizeof(int) * plocal.size()) && h2d(Gp->inbt, inbt.data(), sizeof(int) * n) &&
         h2d(Gp->ifoff, ifoff.data(), sizeof(int) * n) && h2d(Gp->inf, inf.data(), sizeof(int) * n) &&
         h2d(Gp->rinst, rinst.data(), sizeof(int) * rinst.size()) && h2d(Gp->rloc, rloc.data(), sizeof(int) * rloc.size()) &&
         h2d(Gp->aoff, aoff.data(), sizeof(long) * n) && h2d(Gp->ipar, ip.data(), sizeof(InstPar) * n) &&
         h2d(Gp->fargs, Gp->fargs_h.data(), sizeof(ForceArgs) * n) &&
         h2d(Gp->sc_cell, cell.data(), sizeof(int) * cell.size()) && h2d(Gp->sc_off, off.data(), sizeof(int) * off.size()) &&
         h2d(Gp->sc_f, ef.data(), sizeof(int) * ef.size()) && h2d(Gp->sc_w, ew.data(), sizeof(double) * ew.size());
    if (!ok || hipStreamSynchronize(s) != hipSuccess) return bail(IBPM_E_HIP, "building the group failed");
    // scratch that the scatter and rot! leave untouched outside the E' support stays zero
    launch_fill(Gp->qreg, (long)n * nq, 0.0, s);
    launch_fill(Gp->Gt, (long)fG, 0.0, s);
    if (hipStreamSynchronize(s) != hipSuccess) return bail(IBPM_E_HIP, "building the group failed");
    // the members now live in the group's arrays
    for (int i = 0; i < n; i++) {
        ibpm_prob* P = probs[i];
        const long io = (long)i * L;
        Gp->own.push_back({P->G, P->Gs, P->psi, P->nl[0], P->nl[1], P->q, P->q0, P->Fb, P->nl1, P->nl2});
        graphs_clear(P);
        P->G = Gp->G + io * nG; P->Gs = Gp->Gs + io * nG; P->psi = Gp->psi + io * nG;
        P->nl[0] = Gp->nl[0] + io * nG; P->nl[1] = Gp->nl[1] + io * nG; P->nl1 = 0; P->nl2 = 1;
        P->q = Gp->q + io * nq; P->q0 = Gp->q0 + io * nq; P->Fb = Gp->Fb + ifoff[i];
        P->group = Gp;
    }
    *out = Gp;
    return IBPM_OK;
}

int ibpm_group_destroy(ibpm_group* Gp)
{
    if (!Gp) return IBPM_OK;
    (void)hipStreamSynchronize(Gp->stream);
    (void)hipDeviceSynchronize();  // a joint graph of several groups may still run on another group's stream
    group_graphs_clear(Gp);
    multi_drop(Gp);
    const long nG = Gp->g.nG, nq = Gp->g.nq;
    int rc = IBPM_OK;
    for (int i = 0; i < Gp->n; i++) {
        ibpm_prob* P = Gp->probs[i];
        const ibpm_group::Own& o = Gp->own[i];
        const long L = P->mg;  // the member's own levels
        // current state back into the member's own buffers, same roles
        const bool ok =
            hipMemcpy(o.G, P->G, sizeof(double) * L * nG, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.Gs, P->Gs, sizeof(double) * L * nG, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.psi, P->psi, sizeof(double) * L * nG, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.nl0, P->nl[0], sizeof(double) * L * nG, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.nl1b, P->nl[1], sizeof(double) * L * nG, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.q, P->q, sizeof(double) * L * nq, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.q0, P->q0, sizeof(double) * L * nq, hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(o.Fb, P->Fb, sizeof(double) * P->nf, hipMemcpyDeviceToDevice) == hipSuccess;
        if (!ok) rc = IBPM_E_HIP;
        P->G = o.G; P->Gs = o.Gs; P->psi = o.psi; P->nl[0] = o.nl0; P->nl[1] = o.nl1b; P->q = o.q; P->q0 = o.q0;
        P->Fb = o.Fb;
        P->q_stale = Gp->q_stale;  // formed from ψ where it is read, as after the member's own steps
        P->group = nullptr;
    }
    (void)hipStreamSynchronize(Gp->stream);
    Gp->hb.release();
    for (void* p : Gp->allocs) (void)hipFree(p);
    if (Gp->stream) (void)hipStreamDestroy(Gp->stream);
    ibpm_ctx* ctx = Gp->ctx;
    delete Gp;
    return rc == IBPM_OK ? IBPM_OK : fail(ctx, rc, "copying a member's state back failed");
}

int ibpm_group_set_option(ibpm_group* Gp, int option, int value)
{
    if (!Gp) return IBPM_E_ARG;
    // the members' checks hold for the shared grid
    int rc = ibpm_set_option(Gp->probs[0], option, value);
    if (rc != IBPM_OK) return rc;
    if (Gp->pending > 0) return fail(Gp->ctx, IBPM_E_ARG, "wait for the group's queued steps first");
    group_graphs_clear(Gp);
    multi_drop(Gp);
    if (option == IBPM_OPT_YTILE_BATCH) Gp->ytile = value;
    else if (option == IBPM_OPT_YTILE_SINGLE) Gp->ytile_single = value;
    else if (option == IBPM_OPT_XWAVES_BATCH) Gp->xwaves = value;
    else if (option == IBPM_OPT_GRAPHS) Gp->graphs = value == 1;
    else if (option == IBPM_OPT_PIPE) Gp->pipe = value;
    else if (option == IBPM_OPT_YSPLIT) Gp->ysplit = value;
    else if (option == IBPM_OPT_XREUSE) Gp->xreuse = value;
    return IBPM_OK;
}

int ibpm_group_advance_async(ibpm_group* Gp, int nsteps)
{
    if (!Gp || nsteps < 0) return IBPM_E_ARG;
    const size_t want = (size_t)Gp->pending + (size_t)nsteps;
    if (want > Gp->hist_steps) {
        if (Gp->pending > 0)
            return fail(Gp->ctx, IBPM_E_ARG, "history full: call ibpm_group_wait before enqueuing more steps");
        int rc = group_hist(Gp, want);
        if (rc != IBPM_OK) return rc;
    }
    if (nsteps > 0)
        for (int i = 0; i < Gp->n; i++) launch_set_int(Gp->slots + i, Gp->pending, Gp->stream);
    for (int k = 0; k < nsteps; k++) group_step_dev(Gp);
    Gp->pending += nsteps;
    HIPTRY(Gp->ctx, hipGetLastError());
    return IBPM_OK;
}

int ibpm_groups_advance_async(int ngroups, ibpm_group* const* groups, int nsteps)
{
    if (ngroups < 1 || !groups || nsteps < 0) return IBPM_E_ARG;
    for (int a = 0; a < ngroups; a++) {
        if (!groups[a]) return IBPM_E_ARG;
        for (int b = 0; b < a; b++)
            if (groups[b] == groups[a]) return fail(groups[0]->ctx, IBPM_E_ARG, "a group is listed twice");
        if (groups[a]->ctx->device != groups[0]->ctx->device)
            return fail(groups[0]->ctx, IBPM_E_ARG, "the groups must be on one device");
    }
    ibpm_ctx* C = groups[0]->ctx;
    hipStream_t s0 = groups[0]->stream;
    // histories and the per-instance history slots, on each group's stream
    for (int a = 0; a < ngroups; a++) {
        ibpm_group* Gp = groups[a];
        const size_t want = (size_t)Gp->pending + (size_t)nsteps;
        if (want > Gp->hist_steps) {
            if (Gp->pending > 0)
                return fail(C, IBPM_E_ARG, "history full: call ibpm_group_wait before enqueuing more steps");
            int rc = group_hist(Gp, want);
            if (rc != IBPM_OK) return rc;
        }
    }
    if (int rc = multi_events(2 * (size_t)ngroups + 1)) return fail(C, rc, "hipEventCreate failed");
    if (nsteps == 0) return IBPM_OK;
    for (int a = 0; a < ngroups; a++) {
        ibpm_group* Gp = groups[a];
        for (int i = 0; i < Gp->n; i++) launch_set_int(Gp->slots + i, Gp->pending, Gp->stream);
        if (a > 0) {
            HIPTRY(C, hipEventRecord(g_multi_ev[a], Gp->stream));
            HIPTRY(C, hipStreamWaitEvent(s0, g_multi_ev[a], 0));
        }
    }
    for (int k = 0; k < nsteps; k++) {
        // the graph for the groups' current roles, else capture one
        MultiGraph* mg = nullptr;
        for (auto& m : g_multi) {
            if ((int)m.groups.size() != ngroups) continue;
            bool same = true;
            for (int a = 0; a < ngroups && same; a++) {
                ibpm_group* Gp = groups[a];
                const int out = (Gp->nl1 != Gp->nl2) ? Gp->nl1 : 1 - Gp->nl2;
                same = m.groups[a] == Gp && m.G[a] == Gp->G && m.out[a] == out && m.nl2[a] == Gp->nl2 &&
                       m.qs[a] == (int)Gp->q_stale;
            }
            if (same) { mg = &m; break; }
        }
        if (mg) {
            HIPTRY(C, hipGraphLaunch(mg->exec, s0));
            for (int a = 0; a < ngroups; a++) {  // the host side of group_step
                ibpm_group* Gp = groups[a];
                const int out = (Gp->nl1 != Gp->nl2) ? Gp->nl1 : 1 - Gp->nl2;
                Gp->nl1 = Gp->nl2 = out;
                std::swap(Gp->G, Gp->Gs);
                group_roles(Gp);
            }
            continue;
        }
        MultiGraph m;
        for (int a = 0; a < ngroups; a++) {
            ibpm_group* Gp = groups[a];
            m.groups.push_back(Gp);
            m.G.push_back(Gp->G);
            m.out.push_back((Gp->nl1 != Gp->nl2) ? Gp->nl1 : 1 - Gp->nl2);
            m.nl2.push_back(Gp->nl2);
            m.qs.push_back((int)Gp->q_stale);
        }
        hipGraph_t graph = nullptr;
        HIPTRY(C, hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
        bool ok = hipEventRecord(g_multi_ev[0], s0) == hipSuccess;
        for (int a = 1; a < ngroups && ok; a++) ok = hipStreamWaitEvent(groups[a]->stream, g_multi_ev[0], 0) == hipSuccess;
        if (ok)
            for (int a = 0; a < ngroups; a++) group_step(groups[a]);
        for (int a = 1; a < ngroups && ok; a++) {
            ok = hipEventRecord(g_multi_ev[ngroups + a], groups[a]->stream) == hipSuccess &&
                 hipStreamWaitEvent(s0, g_multi_ev[ngroups + a], 0) == hipSuccess;
        }
        const bool ended = hipStreamEndCapture(s0, &graph) == hipSuccess && graph;
        ok = ok && ended && hipGraphInstantiate(&m.exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
            (void)hipGetLastError();
            return fail(C, IBPM_E_HIP, "capturing the joint step of the groups failed");
        }
        HIPTRY(C, hipGraphLaunch(m.exec, s0));
        g_multi.push_back(m);
    }
    // every group's own stream waits for the joint steps (ibpm_group_wait drains it)
    HIPTRY(C, hipEventRecord(g_multi_ev[2 * ngroups], s0));
    for (int a = 0; a < ngroups; a++) {
        if (a > 0) HIPTRY(C, hipStreamWaitEvent(groups[a]->stream, g_multi_ev[2 * ngroups], 0));
        groups[a]->pending += nsteps;
    }
    HIPTRY(C, hipGetLastError());
    return IBPM_OK;
}

int ibpm_group_wait(ibpm_group* Gp, double* cd_hist, double* cl_hist, int32_t* steps_done)
{
    if (!Gp) return IBPM_E_ARG;
    const size_t m = (size_t)Gp->pending;
    std::vector<double> h;
    size_t tot = 0;
    if (m > 0 && Gp->hist) {
        for (int i = 0; i < Gp->n; i++) tot = std::max(tot, Gp->hoff[i] + 2 * Gp->hist_steps * (size_t)Gp->nbh[i]);
        HIPTRY(Gp->ctx, hipMemcpyAsync(Gp->hb.h, Gp->hist, sizeof(double) * tot, hipMemcpyDeviceToHost, Gp->stream));
    }
    HIPTRY(Gp->ctx, hipStreamSynchronize(Gp->stream));
    if (tot) h.assign(Gp->hb.h, Gp->hb.h + tot);
    Gp->pending = 0;
    bool nonfinite = false;
    size_t o = 0;
    for (int i = 0; i < Gp->n; i++) {
        const ibpm_prob* P = Gp->probs[i];
        const size_t nb = (size_t)Gp->nbh[i];
        const double* cd = m ? h.data() + Gp->hoff[i] : nullptr;
        const double* cl = m ? cd + Gp->hist_steps * nb : nullptr;
        int done = (int)m;
        for (size_t k = 0; k < m && done == (int)m; k++)
            for (int b = 0; b < P->nbodies; b++)
                if (!std::isfinite(cl[k * nb + b])) { done = (int)k + 1; break; }
        if (done < (int)m) nonfinite = true;
        if (steps_done) steps_done[i] = done;
        if (cd_hist && m) std::copy(cd, cd + m * nb, cd_hist + o);
        if (cl_hist && m) std::copy(cl, cl + m * nb, cl_hist + o);
        o += m * nb;
    }
    return nonfinite ? fail(Gp->ctx, IBPM_E_NONFINITE, "non-finite CL in a group member") : IBPM_OK;
}

int ibpm_compute_cfl(ibpm_prob* P, double* cfl)
{
    if (!P || !cfl) return IBPM_E_ARG;
    ensure_q(P);
    launch_absmax(P->q, P->g.nq, P->scal, P->stream);
    double m = 0.0;
    HIPTRY(P->ctx, hipMemcpyAsync(&m, P->scal, sizeof(double), hipMemcpyDeviceToHost, P->stream));
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    *cfl = m * P->dt / P->g.h;
    return IBPM_OK;
}

// ---- operator-level entry points --------------------------------------------
struct HostIO {
    ibpm_prob* P;
    std::vector<double*> tmp;
    ~HostIO() { for (double* p : tmp) (void)hipFree(p); }
    double* put(const double* h, size_t n)
    {
        double* d = nullptr;
        if (hipMalloc(&d, sizeof(double) * (n ? n : 1)) != hipSuccess) return nullptr;
        tmp.push_back(d);
        if (h && n) (void)hipMemcpy(d, h, sizeof(double) * n, hipMemcpyHostToDevice);
        else (void)hipMemset(d, 0, sizeof(double) * (n ? n : 1));
        return d;
    }
};

static int get_back(ibpm_prob* P, double* h, const double* d, size_t n)
{
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    HIPTRY(P->ctx, hipMemcpy(h, d, sizeof(double) * n, hipMemcpyDeviceToHost));
    return IBPM_OK;
}

int ibpm_op_lapinv(ibpm_prob* P, int kind, int lev, const double* b, double* x)
{
    if (!P || !b || !x || kind < 0 || kind > 2 || lev < 1 || lev > P->mg) return IBPM_E_ARG;
    HostIO io{P};
    const size_t n = P->g.nG;
    double* db = io.put(b, n);
    double* dx = io.put(nullptr, n);
    if (!db || !dx) return fail(P->ctx, IBPM_E_HIP, "alloc");
    if (P->prec >= 0 && kind != 1) {
        // the mixed-precision transforms (A is applied by its stencil in that form)
        MxResidArgs r = mx_rargs(P, MX_RES_COPY, lev - 1);
        r.G = db;
        launch_mx_resid(r, 1, P->stream);
        mx_apply(P, kind == 0 ? MX_OP_DINV : MX_OP_AINV, lev - 1, 1, MX_EPI_ADD, dx, nullptr);
        return get_back(P, x, dx, n);
    }
    lap_level(P, kind == 0 ? Y_DINV : (kind == 1 ? Y_A : Y_AINV), lev - 1, db, dx, false);
    return get_back(P, x, dx, n);
}

int ibpm_op_lapinv_noalloc(ibpm_prob* P, int kind, int lev, const double* b, double* x, size_t n)
{
    if (!P || !b || !x || kind < 0 || kind > 2 || lev < 1 || lev > P->mg) return IBPM_E_ARG;
    if (n != (size_t)P->g.nG) return fail(P->ctx, IBPM_E_ARG, "n must be (nx-1)(ny-1) of the problem's grid");
    if (n > (size_t)IBPM_NOALLOC_MAX_POINTS)
        return fail(P->ctx, IBPM_E_ARG, "the level has more points than the fixed scratch holds (IBPM_NOALLOC_MAX_POINTS)");
    if (!P->ctx->scratch) return fail(P->ctx, IBPM_E_HIP, "no scratch address for this device");
    // the x-pass writes every entry of the scratch before the y-pass reads it
    lap_level(P, kind == 0 ? Y_DINV : (kind == 1 ? Y_A : Y_AINV), lev - 1, b, x, false, 1, 1 << 30,
              P->ctx->scratch);
    HIPTRY(P->ctx, hipGetLastError());
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    return IBPM_OK;
}

int ibpm_op_dst2(ibpm_prob* P, const double* x, double* y)
{
    if (!P || !x || !y) return IBPM_E_ARG;
    // 2D RODFT00 from two x-direction passes: transform along x, transpose on
    // the host, transform the former y lines with the y plan, transpose back.
    HostIO io{P};
    const int n1 = P->g.n1, n2 = P->g.n2;
    const size_t n = P->g.nG;
    double* dx = io.put(x, n);
    double* dy = io.put(nullptr, n);
    if (!dx || !dy) return fail(P->ctx, IBPM_E_HIP, "alloc");
    XArgs a = xargs(P);
    a.src = dx; a.dst = dy; a.nlev = 1; a.lev0 = 0;
    launch_xpass(X_PLAIN, a, P->stream);
    std::vector<double> hx(n), ht(n);
    int rc = get_back(P, hx.data(), dy, n);
    if (rc) return rc;
    for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) ht[(size_t)j + (size_t)n2 * i] = hx[(size_t)i + (size_t)n1 * j];
    // transform along the former y (now contiguous) with the y plan
    GridDev gt = P->g;
    std::swap(gt.n1, gt.n2);
    XArgs b = xargs(P);
    b.g = gt;
    b.zhi = gt.n2;
    b.pl = P->ty.plan;
    double* dt2 = io.put(ht.data(), n);
    double* dt3 = io.put(nullptr, n);
    b.src = dt2; b.dst = dt3; b.nlev = 1; b.lev0 = 0;
    launch_xpass(X_PLAIN, b, P->stream);
    rc = get_back(P, ht.data(), dt3, n);
    if (rc) return rc;
    for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) y[(size_t)i + (size_t)n1 * j] = ht[(size_t)j + (size_t)n2 * i];
    return IBPM_OK;
}

int ibpm_op_nonlinear(ibpm_prob* P, int lev, const double* q, const double* q0, const double* G, double* out)
{
    if (!P || !q || !q0 || !G || !out || lev < 1 || lev > P->mg) return IBPM_E_ARG;
    HostIO io{P};
    const size_t nq = P->g.nq * P->mg, nG = P->g.nG * P->mg;
    double *dq = io.put(q, nq), *dq0 = io.put(q0, nq), *dG = io.put(G, nG), *dout = io.put(nullptr, P->g.nG);
    if (!dq || !dq0 || !dG || !dout) return fail(P->ctx, IBPM_E_HIP, "alloc");
    NLArgs na{};
    na.q = dq; na.q0 = dq0; na.G = dG; na.out = dout; na.nlev = 1; na.lev0 = lev - 1; na.mg = P->mg; na.g = P->g;
    na.QB = P->lin ? P->QB : nullptr;  // the given q0 is streamed (q0_uniform stays 0)
    na.GB = P->lin ? P->GB : nullptr;
    launch_nonlinear(na, P->stream);
    return get_back(P, out, dout, P->g.nG);
}

int ibpm_op_vort2flux(ibpm_prob* P, double* G, double* psi, double* q)
{
    if (!P || !G || !psi || !q) return IBPM_E_ARG;
    HostIO io{P};
    const size_t nq = P->g.nq * P->mg, nG = P->g.nG * P->mg;
    double *dG = io.put(G, nG), *dpsi = io.put(nullptr, nG), *dq = io.put(nullptr, nq);
    if (!dG || !dpsi || !dq) return fail(P->ctx, IBPM_E_HIP, "alloc");
    vort2flux_dev(P, dG, dpsi, dq, CURL_ALL);
    int rc = get_back(P, G, dG, nG);
    if (!rc) rc = get_back(P, psi, dpsi, nG);
    if (!rc) rc = get_back(P, q, dq, nq);
    return rc;
}

int ibpm_op_regT(ibpm_prob* P, const double* q, double* fb)
{
    if (!P || !q || !fb) return IBPM_E_ARG;
    HostIO io{P};
    double *dq = io.put(q, P->g.nq), *dfb = io.put(nullptr, P->nf);
    if (!dq || !dfb) return fail(P->ctx, IBPM_E_HIP, "alloc");
    GatherArgs ga{};
    ga.q = dq; ga.q0 = nullptr; ga.w = P->w; ga.bidx = P->bidx; ga.out = dfb; ga.nbt = P->nbt; ga.g = P->g;
    launch_gather(ga, P->stream);
    return get_back(P, fb, dfb, P->nf);
}

int ibpm_op_reg(ibpm_prob* P, const double* fb, double* q)
{
    if (!P || !fb || !q) return IBPM_E_ARG;
    HostIO io{P};
    double *dfb = io.put(fb, P->nf), *dq = io.put(nullptr, P->g.nq);
    if (!dq || !dfb) return fail(P->ctx, IBPM_E_HIP, "alloc");
    ScatterArgs sa{};
    sa.cell = P->sc_cell; sa.off = P->sc_off; sa.ent_f = P->sc_f; sa.ent_w = P->sc_w;
    sa.fb = dfb; sa.q = dq; sa.ncell = P->sc_ncell;
    launch_scatter(sa, P->stream);
    return get_back(P, q, dq, P->g.nq);
}

int ibpm_op_rot(ibpm_prob* P, const double* q, double* G)
{
    if (!P || !q || !G) return IBPM_E_ARG;
    HostIO io{P};
    double *dq = io.put(q, P->g.nq), *dG = io.put(nullptr, P->g.nG);
    if (!dq || !dG) return fail(P->ctx, IBPM_E_HIP, "alloc");
    launch_rot(dq, dG, 2, P->g.nx, 2, P->g.ny, P->g, P->stream);
    return get_back(P, G, dG, P->g.nG);
}

int ibpm_set_option(ibpm_prob* P, int option, int value)
{
    if (!P) return IBPM_E_ARG;
    if (int rc = check_idle(P)) return rc;
    graphs_clear(P);
    if (option == IBPM_OPT_XWAVES_BATCH) {
        if (value != 4 && value != 8 && value != 12 && value != 14)
            return fail(P->ctx, IBPM_E_ARG, "x-pass work-group lines must be 4, 8, 12 or 14");
        if (xpass_lds_bytes(P->tx.plan, value) > 160 * 1024)
            return fail(P->ctx, IBPM_E_ARG, "x-pass work-group does not fit in LDS for this nx");
        P->xwaves = value;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_YTILE_BATCH || option == IBPM_OPT_YTILE_SINGLE) {
        if (value != 4 && value != 8 && value != 12) return fail(P->ctx, IBPM_E_ARG, "y tile must be 4, 8 or 12");
        if (ypass_lds_bytes(P->ty.plan, P->g.n2, value) > 160 * 1024)
            return fail(P->ctx, IBPM_E_ARG, "y tile does not fit in LDS for this ny");
        (option == IBPM_OPT_YTILE_BATCH ? P->ytile : P->ytile_single) = value;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_PRECISION) {
        if (value < 0 || value > 2) return fail(P->ctx, IBPM_E_ARG, "precision option is 0 (FP64), 1 (FP32-MFMA) or 2 (BF16x3-MFMA)");
        if (value > 0 && P->motion == IBPM_MOTION_FUNCTION)
            return fail(P->ctx, IBPM_E_UNSUPPORTED, "the mixed-precision form runs static and MovingGrid bodies");
        if (value > 0) {
            int rc = mx_setup(P);
            if (rc != IBPM_OK) return rc;
        }
        P->prec = value == 0 ? -1 : (value == 1 ? MX_FP32 : MX_BF16X3);
        return IBPM_OK;
    }
    if (option == IBPM_OPT_MX_SWEEPS) {
        if (value != 1 && value != 2) return fail(P->ctx, IBPM_E_ARG, "mixed-precision sweeps are 1 or 2");
        P->mx_sweeps = value;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_GRAPHS) {
        if (value != 0 && value != 1) return fail(P->ctx, IBPM_E_ARG, "graphs option is 0 or 1");
        P->graphs = value == 1;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_PIPE) {
        if (value != 0 && value != 1) return fail(P->ctx, IBPM_E_ARG, "pipe option is 0 or 1");
        P->pipe = value;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_YSPLIT) {
        if (value != 0 && value != 1) return fail(P->ctx, IBPM_E_ARG, "ysplit option is 0 or 1");
        P->ysplit = value;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_XREUSE) {
        if (value != 0 && value != 1) return fail(P->ctx, IBPM_E_ARG, "xreuse option is 0 or 1");
        P->xreuse = value;
        return IBPM_OK;
    }
    if (option == IBPM_OPT_PROFILE) {
        if (value < 0 || value > 2) return fail(P->ctx, IBPM_E_ARG, "profile option is 0, 1 or 2");
        HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
        collect_profile(P);
        P->prof = value;
        P->prof_ms = 0.0;
        P->prof_n = 0;
        return IBPM_OK;
    }
    return fail(P->ctx, IBPM_E_ARG, "unknown option");
}

int ibpm_get_option(ibpm_prob* P, int option, int* value)
{
    if (!P || !value) return IBPM_E_ARG;
    if (option == IBPM_OPT_YTILE_BATCH) *value = P->ytile;
    else if (option == IBPM_OPT_YTILE_SINGLE) *value = P->ytile_single;
    else if (option == IBPM_OPT_PROFILE) *value = P->prof;
    else if (option == IBPM_OPT_GRAPHS) *value = P->graphs ? 1 : 0;
    else if (option == IBPM_OPT_XWAVES_BATCH) *value = P->xwaves;
    else if (option == IBPM_OPT_PIPE) *value = P->pipe;
    else if (option == IBPM_OPT_YSPLIT) *value = P->ysplit;
    else if (option == IBPM_OPT_XREUSE) *value = P->xreuse;
    else if (option == IBPM_OPT_PRECISION) *value = P->prec < 0 ? 0 : (P->prec == MX_FP32 ? 1 : 2);
    else if (option == IBPM_OPT_MX_SWEEPS) *value = P->mx_sweeps;
    else return fail(P->ctx, IBPM_E_ARG, "unknown option");
    return IBPM_OK;
}

int ibpm_profile_stats(ibpm_prob* P, double* avg_ms, int64_t* count)
{
    if (!P || !avg_ms || !count) return IBPM_E_ARG;
    HIPTRY(P->ctx, hipStreamSynchronize(P->stream));
    collect_profile(P);
    *avg_ms = P->prof_n ? P->prof_ms / (double)P->prof_n : 0.0;
    *count = P->prof_n;
    return IBPM_OK;
}

int ibpm_time_kernel(ibpm_prob* P, int which, int reps, double* ms_avg)
{
    if (!P || !ms_avg || reps < 1) return IBPM_E_ARG;
    hipEvent_t e0, e1;
    HIPTRY(P->ctx, hipEventCreate(&e0));
    HIPTRY(P->ctx, hipEventCreate(&e1));
    const int L = P->mg;
    auto one = [&]() {
        if (which == 0) {
            XArgs a = xargs(P);
            a.src = P->G; a.dst = P->T1; a.nlev = L; a.lev0 = 0;
            launch_xpass(X_PLAIN, a, P->stream);
        } else if (which == 1) {
            launch_ypass(yargs(P, Y_A, 0, L, P->T1), P->ytile, P->stream);
        } else if (which == 2) {
            NLArgs na{};
            na.q = P->q; na.q0 = P->q0; na.G = P->G; na.out = P->T2; na.nlev = L; na.lev0 = 0; na.mg = L; na.g = P->g;
            set_nl_q0(P, na);
            launch_nonlinear(na, P->stream);
        } else if (which == 4) {
            XArgs a = xargs(P);
            a.src = P->G; a.dst = P->T1; a.nlev = 1; a.lev0 = 0;
            launch_xpass(X_PLAIN, a, P->stream);
        } else if (which == 5) {
            launch_ypass(yargs(P, Y_DINV, 0, 1, P->T1), P->ytile_single, P->stream);
        } else if (which == 6) {
            CurlArgs ca{};
            ca.psi = P->psi; ca.q = P->T2; ca.qstride = 0; ca.nlev = 1; ca.lev0 = 0; ca.mg = L; ca.g = P->g;
            launch_curl(ca, P->stream);
        } else if (which == 7) {
            launch_copy(P->T2, P->T1, P->g.nG * L, P->stream);
        } else if (which == 8 && L > 1) {
            YArgs y = yargs(P, Y_DINV, 0, 1, P->T1);
            y.bcr = P->psi + P->g.nG;
            y.edge = P->edge;
            launch_ypass(y, P->ytile_single, P->stream);
        } else if (which >= 9 && which <= 11) {
            // mixed-precision sine-matrix transforms: 9 x over all levels, 10 y
            // with the eigenvalue scaling over all levels, 11 x on level 1
            MxGemmArgs a{};
            a.istride = a.ostride = (long)P->P1 * P->P2;
            a.fstride = P->g.nG;
            a.P1 = P->P1; a.P2 = P->P2;
            a.op = MX_OP_AINV; a.cx = P->cx; a.cy = P->cy; a.dt = P->dt; a.Re = P->Re;
            a.scale = 1.0 / (double)(4L * P->g.nx * P->g.ny);
            a.lscale = P->mx_lscale; a.g = P->g;
            a.part = P->mx_part; a.partcap = (long)MX_SPLIT_MAX * P->P1 * P->P2 * P->mg;
            a.in = P->mx_buf[0]; a.out = P->mx_buf[1];
            if (which == 10) { a.sf = P->mx_sf2; a.sh = P->mx_sh2; a.sl = P->mx_sl2; }
            else { a.sf = P->mx_sf1; a.sh = P->mx_sh1; a.sl = P->mx_sl1; }
            launch_mx_gemm(P->prec, which == 10 ? 1 : 0, which == 10 ? MX_EPI_SCALE : MX_EPI_STORE, a,
                           which == 11 ? 1 : L, P->stream);
        }
    };
    if ((which >= 9 && which <= 11) && P->prec < 0) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return fail(P->ctx, IBPM_E_ARG, "the mixed-precision kernels need IBPM_OPT_PRECISION set");
    }
    if (which < 0 || which > 11 || which == 3 || (which == 8 && L < 2)) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return fail(P->ctx, IBPM_E_ARG, "unknown kernel id (whole steps are timed through ibpm_advance_n)");
    }
    one();
    HIPTRY(P->ctx, hipEventRecord(e0, P->stream));
    for (int r = 0; r < reps; r++) one();
    HIPTRY(P->ctx, hipEventRecord(e1, P->stream));
    HIPTRY(P->ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPTRY(P->ctx, hipEventElapsedTime(&ms, e0, e1));
    *ms_avg = (double)ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return IBPM_OK;
}

}  // extern "C"
