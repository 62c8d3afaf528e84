// kernels.h -- argument blocks and launchers of the HIP kernels (kernels.hip).
#pragma once
#include "ibpm_dev.h"

namespace ibpm {

constexpr int MAXLEV = 16;

// Per-instance parameters of a group of same-shape problems stepped together.
// A batch over all levels of all instances has entries e = inst * L + lev
// (levmod = L); its kernels take dt, Re, the AB2 coefficients and the uniform
// base flux of entry e from ipar[e / L].
struct InstPar {
    double c1, c2;      // β1 dt, β2 dt
    double dt, Re;
    double q0vx, q0vy;  // inertial base flux per unit hc
    int L;              // levels of the instance (a group keeps every instance at its largest count;
    int pad;            //   entries of levels an instance does not have are skipped)
};

// ---- DST passes -----------------------------------------------------------
// X_ROT: the input line is rot! (curl.jl:22-34) of the level-1 flux field src
// (the E' output of project_circ!), formed in the pass
enum XMode { X_PLAIN = 0, X_ROT = 1, X_SUB = 2, X_TRIAL = 3 };

struct XArgs {
    const double* src;
    double* dst;
    long lstride;         // element stride between batch entries (levels)
    int nlev;             // batch size
    int lev0;             // 0-based grid level of batch entry 0
    int mg;
    // X_PLAIN: when edge != null, line j (1-based) also stores its values at
    // i = nx/4 and 3nx/4 to edge[2 (j - 1)] and edge[2 (j - 1) + 1] (entry b at
    // edge + b * edstride): the columns get_bc! of the next finer level samples
    double* edge;
    long edstride;
    // X_TRIAL: src = Γ (all levels); rhs = ((A Γ + c1 N1) + c2 N0) + rhsbc
    // with A Γ from the 5-point stencil, then the x-DST of rhs
    const double* n1;
    const double* n0;
    const double* gam;    // state Γ (all levels) for rhsbc of level lev from lev+1
    double c1, c2;
    double dt, Re;        // rhsbc factor 0.25 dt / (Re hc^2) is formed per level in the kernel
    // X_PLAIN: input lines j (1-based) outside [zlo, zhi] are known to be zero;
    // their output lines are written as zeros without a transform
    int zlo, zhi;
    // X_PLAIN, lsel != 0: only the lines llo..lhi (0-based) are transformed
    // (the launch covers no others); keep != 0: of those, only lines keep0 and
    // keep1 store their output (the edge columns are stored for every line)
    int lsel, llo, lhi;
    int keep, keep0, keep1;
    // X_PLAIN over a batch (nlev > 1): waves (lines) per work-group, 8, 12 or
    // 14; anything else (0) gives the default 4
    int waves;
    // X_TRIAL over a group: entry e = lev0 + b is level e % levmod of
    // instance e / levmod, whose c1, c2, dt, Re come from ipar (0 / null: one problem);
    // fixlev: every entry is at level lev0 (a batch of fields, or instance b of a group)
    int levmod;
    int fixlev;
    const InstPar* ipar;
    // entries of a level below minlev (0-based) are skipped
    int minlev;
    DstPlan pl;
    GridDev g;
};
void launch_xpass(int mode, const XArgs& a, hipStream_t s);
// LDS bytes of an x-pass work-group of `waves` lines
size_t xpass_lds_bytes(const DstPlan& pl, int waves);

enum YOp { Y_DINV = 0, Y_A = 1, Y_AINV = 2 };
struct YArgs {
    double* buf;
    long lstride;
    int nlev, lev0;
    int fixlev;           // 1: every batch entry is level lev0 (a batch of fields); 0: entry b is level lev0 + b
    int op;
    const double* cx;   // cos(pi i / nx), i = 1..nx-1
    const double* cy;   // cos(pi j / ny), j = 1..ny-1
    double dt;
    double Re;            // divisor (2 Re) hc^2 is formed per level in the kernel
    double scale;         // 1 / (4 nx ny)
    // Y_DINV with boundary terms: the x-transform of apply_bc!(get_bc!(ψ of
    // the next coarser level), bcfac), added to the x-transformed lines before
    // the first y-transform (entries whose level is the instance's coarsest
    // have none).  bcr: that coarser level (entry b at bcr + b * bcstride);
    // edge: its columns nx/4, 3nx/4 as written by the X_PLAIN edge output
    // (entry b at edge + b * edstride); sinx: sin(pi m / nx), m = 0..nx/2.
    // bcr == null: no boundary terms.
    const double* bcr;
    const double* edge;
    long bcstride, edstride;
    double bcfac;
    const double* sinx;
    // over a group: entry e = lev0 + b is level e % levmod of instance
    // e / levmod (fixlev: entry b is instance b at level lev0); dt and Re of
    // the instance from ipar (0 / null: one problem)
    int levmod;
    const InstPar* ipar;
    // rsel != 0: input rows (0-based lines) outside rlo..rhi are known to be
    // zero; they are not read.  wsel != 0: only output rows wlo..whi are
    // stored (the others are not read after the pass)
    int rsel, rlo, rhi;
    int wsel, wlo, whi;
    // pipe != 0: a launch of many tiles (no boundary terms, 4-column tiles)
    // runs as persistent work-groups that prefetch their next tile
    int pipe;
    // ysrc != null: the input rows outside clo..chi (0-based) are not read
    // from buf but are afac * ysrc (ysub null) or afac * (ysrc - ysub), entry b
    // at ysrc / ysub + b * lstride: rows whose x-transform is known as a
    // multiple of another field's (the DST-I is its own inverse up to 2 nx)
    const double* ysrc;
    const double* ysub;
    int clo, chi;
    double afac;
    // split != 0 (1023-point columns, 4-column tiles, launches of at most one
    // tile per CU: a single level): each column is transformed by two
    // wavefronts (dst1024_split); 0: one wavefront per column
    int split;
    DstPlan pl;
    GridDev g;
};
void launch_ypass(const YArgs& a, int tile, hipStream_t s);
// LDS bytes one y-pass workgroup needs (tables, cos table, tile slots)
size_t ypass_lds_bytes(const DstPlan& pl, int n2, int tile);
// LDS bytes of the split y-pass work-group (ny = 1024), with or without boundary terms
size_t ypass_split_lds_bytes(int ny, bool bc);

// ---- field kernels ---------------------------------------------------------
struct NLArgs {
    const double* q;
    // when non-null, the fluxes are formed from psi (all levels: entry e at
    // e nG, its boundary values from entry e + 1) and q is not read
    const double* psi;
    const double* q0;
    const double* G;
    double* out;
    int nlev, lev0, mg;
    // q0 uniform per level and component (inertial base flux): q0 = (vx, vy) hc,
    // so the kernel need not stream q0
    int q0_uniform;
    double q0vx, q0vy;
    // LinearizedIBProblem (problem-types.jl:47-87): when QB != null the
    // product is Q x ΓB + QB x Γ with Q = avg_flux of q alone (nonlinear.jl:45-65,
    // 144-163); QB (nq per level, averaged) and GB (ΓB, nG per level)
    const double* QB;
    const double* GB;
    // over a group: entry e = lev0 + b is level e % levmod of instance
    // e / levmod, with that instance's base flux from ipar (0 / null: one problem)
    int levmod;
    const InstPar* ipar;
    GridDev g;
};
void launch_nonlinear(const NLArgs& a, hipStream_t s);

struct CurlArgs {
    const double* psi;  // all levels (for the bc of level lev from lev+1)
    double* q;          // output: entry b is level lev0 + b
    long qstride;
    int nlev, lev0, mg;
    // colbatch = 1: the nlev entries are fields at level lev0, field b's level
    // l at psi + l * psl + b * psb
    int colbatch;
    long psl, psb;
    // over a group (colbatch = 0): entry e = lev0 + b is level e % levmod;
    // ipar: the instances' level counts (colbatch: entry b is instance b)
    int levmod;
    const InstPar* ipar;
    GridDev g;
};
void launch_curl(const CurlArgs& a, hipStream_t s);

struct CoarsifyArgs {
    double* Gc;
    const double* G;
    int i0, ic0, ci, j0, jc0, cj;
    int nb;               // batch of fields (0 or 1: one), field b at Gc/G + b * bstride
    long bstride;
    // a group: field b is instance b, which has level clev (1-based, the level
    // written) only if clev <= ipar[b].L
    const InstPar* ipar;
    int clev;
    GridDev g;
};
void launch_coarsify(const CoarsifyArgs& a, hipStream_t s);

// rot! over the rectangle [ib0, ib1] x [jb0, jb1] of 1-based Γ indices (i, j
// here are the flux-grid indices 2..nx, 2..ny of rot!, curl.jl:26)
void launch_rot(const double* q, double* G, int ib0, int ib1, int jb0, int jb1, const GridDev& g, hipStream_t s);

struct GatherArgs {
    const double* q;
    // when psi is non-null (level 1), the fluxes are formed from psi by curl!
    // with the boundary values of psic (the level-2 psi; null: zero) instead
    // of being read from q.  Over a group, instance inst's psi is at
    // psi + inst * psistride, its psic nG further on (when psic != null and
    // the instance has two levels or more, ipar[inst].L)
    const double* psi;
    const double* psic;
    long psistride;
    const InstPar* ipar;
    const double* q0;  // may be null
    const double* w;   // [(2k + c) * 169 + l + 13 m]
    const int* bidx;   // [k] = bx, [nbt + k] = by (1-based)
    double* out;       // 2 nbt
    int nbt;
    int nb;            // batch of fields (0 or 1: one): q + b * qstride -> out + b * ostride
    long qstride, ostride;
    // over a group (pinst != null): nbt is the total of all instances' points;
    // point k belongs to instance pinst[k] as its point plocal[k]; that
    // instance reads q + inst * qstride, q0 + inst * q0stride and writes its
    // F̃b at out + ifoff[inst] (inbt[inst] points)
    const int* pinst;
    const int* plocal;
    const int* inbt;
    const int* ifoff;
    long q0stride;
    GridDev g;
};
void launch_gather(const GatherArgs& a, hipStream_t s);

// E' e_f for f = f0 .. f0 + nb - 1 (unit force vectors, reg! of
// interface-coupling.jl:99-119): field b (at q + b * qstride, zero on entry)
// receives the 13 x 13 weights of body point f mod nbt, component f / nbt
void launch_unit_reg(double* q, long qstride, int nb, int f0, const int* bidx, const double* w, int nbt,
                     const GridDev& g, hipStream_t s);
// rot! over a rectangle for nb fields: q + b * qstride -> G + b * gstride
void launch_rot_batch(const double* q, long qstride, double* G, long gstride, int nb, int ib0, int ib1, int jb0,
                      int jb1, const GridDev& g, hipStream_t s);

struct ScatterArgs {
    const int* cell;   // flux index of each touched cell
    const int* off;    // CSR offsets (ncell + 1)
    const int* ent_f;  // F̃b index of each contribution
    const double* ent_w;
    const double* fb;
    double* q;
    int ncell;
};
void launch_scatter(const ScatterArgs& a, hipStream_t s);
struct ForceArgs;
// the scatter and update_stress! (k_forces) in one launch (both read F̃b only)
void launch_scatter_forces(const ScatterArgs& a, const ForceArgs& fa, hipStream_t s);
// y = A x by the row GEMV of launch_gemv_forces, without the forces
void launch_gemv(const double* AT, const double* x, double* y, int n, hipStream_t s);

struct ForceArgs {
    const double* Fb;
    const double* ds;
    const int* boff;  // body offsets (nbodies + 1)
    double* fb;
    double* CD;
    double* CL;
    double* cd_hist;  // may be null
    double* cl_hist;
    // with slot != null the history entry is [(*slot) * nbh + body] of cd_hist
    // and cl_hist, and the kernel advances *slot by one (graph replays of a
    // step then need no per-step arguments)
    int* slot;
    int nbh;
    int nbodies, nbt;
    double hdt;       // h / dt
};
// y = Binv x (AT = Binv transposed, row-major rows), then update_stress! on y
void launch_gemv_forces(const double* AT, const double* x, double* y, int n, const ForceArgs& fa, hipStream_t s);
// update_stress! alone, on an F̃b already in place (MotionFunction: F̃b comes from the CG solve)
void launch_forces(const ForceArgs& fa, hipStream_t s);
// a group: y_i = Binv_i x_i for every instance (rows of all instances: row r
// is row rloc[r] of instance rinst[r], whose Binv (transposed) starts at AT +
// aoff[inst], has inf[inst] rows and reads/writes x, y at ifoff[inst]), then
// update_stress! per instance (fa: device array of n argument blocks)
void launch_gemv_forces_group(const double* AT, const long* aoff, const double* x, double* y, const int* rinst,
                              const int* rloc, const int* ifoff, const int* inf, int nrows, const ForceArgs* fa,
                              int n, hipStream_t s);

struct BaseFluxArgs {
    double* q0;
    int mg;
    double Ux0, Uy0, Om;   // translation in the body frame and Ω = -θ̇
    double offx, offy;
    GridDev g;
};
void launch_base_flux_moving(const BaseFluxArgs& a, hipStream_t s);
void launch_base_flux_inertial(double* q0, int mg, double vx, double vy, const GridDev& g, hipStream_t s);

void launch_fill(double* p, long n, double v, hipStream_t s);
void launch_set_int(int* p, int v, hipStream_t s);
void launch_copy(double* dst, const double* src, long n, hipStream_t s);
void launch_set_unit(double* p, int n, int j, hipStream_t s);
void launch_absmax(const double* p, long n, double* out, hipStream_t s);  // out: single double (>= 0)

// inv(M) of an n x n matrix by Gauss-Jordan elimination with row pivoting
// (setup.hip): M is the n x 2n column-major augmented [B | I], overwritten,
// the inverse left in its right half; piv, pval, singular: one device int,
// double, int (singular set to 1 on a zero pivot column); f: n doubles
void launch_gauss_jordan(double* M, int n, int* piv, double* pval, int* singular, double* f, hipStream_t s);

void kernels_init();  // sets large-LDS attributes

}  // namespace ibpm
