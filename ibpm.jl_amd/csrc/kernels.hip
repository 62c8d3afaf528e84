// kernels.hip -- HIP kernels of the advance! step for gfx950.
//
// All arithmetic is FP64.  Element-wise expressions that the reference fixes
// (stencils, eigenvalue scaling, right-hand-side assembly) are written in the
// reference's order of operations; the library is compiled with
// -ffp-contract=off so each a*b+c stays two rounded operations as in Julia.
// Transform internals (FFT butterflies) use explicit fma().
#include "kernels.h"

namespace ibpm {

// ===========================================================================
// DST x-pass: one wavefront per x-line (fixed j) of one level.
// ===========================================================================
constexpr int XWAVES = 4;  // wavefronts per workgroup

// Batch entry b of a launch -> (instance, level).  One problem: entry b is
// level lev0 + b (fixlev: level lev0 for every entry, a batch of fields).  A
// group: entry e = lev0 + b is level e % levmod of instance e / levmod
// (fixlev: instance b at level lev0).  `active` is false for a level the
// instance does not have; `L` is its level count.
struct Entry {
    int inst, lev, L;
    bool active;
};
__device__ __forceinline__ Entry batch_entry(int b, int lev0, int levmod, int fixlev, int mg, const InstPar* ip)
{
    Entry r;
    const int e = lev0 + b;
    r.inst = fixlev ? b : (levmod ? e / levmod : 0);
    r.lev = fixlev ? lev0 : (levmod ? e % levmod : e);
    r.L = ip ? ip[r.inst].L : mg;
    r.active = r.lev < r.L;
    return r;
}
constexpr int XPL = (MAXM + WAVE - 1) / WAVE;  // line elements per lane

// Per-length line layout in an LDS slot.  PT = 512 (lines of 1023) uses the
// guarded layout of dst1024_line; every other length stores element n at n.
template <int PT>
__device__ __forceinline__ int lat(int n)
{
    if constexpr (PT == 512) return f1024::gsl(n + 1);
    else return n;
}
template <int PT>
__host__ __device__ inline int slot_doubles_t(const DstPlan& pl)
{
    if constexpr (PT == 512) return f1024::SLOT;
    else return dst_slot_doubles(pl);
}
template <int PT>
__host__ __device__ inline int tab_doubles_t(const DstPlan& pl)
{
    if constexpr (PT == 512) return f1024::TAB;
    else return dst_tab_doubles(pl);
}
template <int PT, int NT>
__device__ __forceinline__ DstTab load_tab_t(double* lds, const DstPlan& pl, int tid)
{
    if constexpr (PT == 512) {
        Tab1024Regs<NT> R;
        tab1024_issue(R, pl, tid);
        return tab1024_store(lds, R, tid);
    } else {
        return load_dst_tab(lds, pl, tid, NT);
    }
}
// line length: 1023 for PT = 512 (a compile-time constant), else the grid's
template <int PT>
__device__ __forceinline__ int line_len(int n)
{
    if constexpr (PT == 512) return f1024::M - 1;
    else return n;
}
// zero guard cells of a slot (guarded layout only)
template <int PT>
__device__ __forceinline__ void set_guards(double* s)
{
    if constexpr (PT == 512) {
        s[f1024::gsl(0)] = 0.0;
        s[f1024::gsl(f1024::M)] = 0.0;
    }
}

// one line transform: compile-time length PT > 0, else the plan's runtime length
template <int PT>
__device__ __forceinline__ void dst_line(double* s, const DstPlan& pl, const DstTab& tb, int lane)
{
    if constexpr (PT == 512) dst1024_line(s, tb, lane, F1024Store{s});
    else if constexpr (PT > 0) dst1_line_t<PT>(s, tb, lane);
    else dst1_line(s, pl, tb, lane);
}

template <int MODE, int PT, int XW>
__global__ void __launch_bounds__(XW * WAVE) k_xpass(XArgs a)
{
    extern __shared__ double lds[];
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const int n1 = line_len<PT>(a.g.n1), n2 = a.g.n2;
    const int lend = a.lsel ? a.lhi + 1 : n2;  // one past the last line of the launch
    const int line = (a.lsel ? a.llo : 0) + blockIdx.x * XW + w;  // 0-based j - 1
    const int b = blockIdx.y;
    const int j = line + 1;
    const Entry en = batch_entry(b, a.lev0, a.levmod, a.fixlev, a.mg, a.ipar);
    // the whole work-group: a level its instance does not have, or one the launch skips
    if (!en.active || en.lev < a.minlev) return;
    const int ntab = tab_doubles_t<PT>(a.pl);
    const DstTab tb = load_tab_t<PT, XW * WAVE>(lds, a.pl, threadIdx.x);
    __syncthreads();
    if (line >= lend) return;
    if constexpr (MODE == X_PLAIN) {
        if (j < a.zlo || j > a.zhi) {  // zero line: its transform is zero
            double* dz = a.dst + b * a.lstride + (long)line * n1;
            for (int n = lane; n < n1; n += WAVE) dz[n] = 0.0;
            if (a.edge && lane < 2) a.edge[b * a.edstride + 2 * line + lane] = 0.0;
            return;
        }
    }
    // every global load of the line is in flight before the first wait
    const double* src = a.src + b * a.lstride + (long)line * n1;
    double v[XPL];
    if constexpr (MODE == X_ROT) {
        // Γ(m, j), m = n + 1: rot! at flux-grid (i, jj) = (m + 1, j + 1) of the
        // x-fluxes (i, jj-1), (i, jj) and y-fluxes (i, jj), (i-1, jj), as k_rot forms it
        const double* q = a.src + b * a.lstride;
        const int nx = a.g.nx;
        const double* qx = q + (long)(nx + 1) * (j - 1);  // x-fluxes of row jj - 1 = j
        const double* qy = q + a.g.nu + (long)nx * j;     // y-fluxes of row jj = j + 1
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            const int n = lane + WAVE * t;
            if (n < n1) {
                const int i = n + 2;
                const double qa = qx[i - 1], qb = qx[(i - 1) + (nx + 1)];
                const double ya = qy[i - 1], yb = qy[i - 2];
                v[t] = (qa - qb) + (ya - yb);
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            int n = lane + WAVE * t;
            if (n < n1) v[t] = src[n];
        }
    }
    // X_TRIAL: the rows j - 1 and j + 1 of Γ (zero outside), N1 and N0 of the
    // line, in flight with the row's own loads
    double vm[XPL], vp[XPL], n1v[XPL], u[XPL];
    if constexpr (MODE == X_TRIAL) {
        const double* n1p = a.n1 + b * a.lstride + (long)line * n1;
        const double* n0p = a.n0 + b * a.lstride + (long)line * n1;
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            const int n = lane + WAVE * t;
            vm[t] = vp[t] = n1v[t] = u[t] = 0.0;
            if (n < n1) {
                if (line > 0) vm[t] = src[n - n1];
                if (line < n2 - 1) vp[t] = src[n + n1];
                n1v[t] = n1p[n];
                u[t] = n0p[n];
            }
        }
    }
    const int ld = slot_doubles_t<PT>(a.pl);
    double* s = lds + ntab + w * ld;
    if (lane == 0) set_guards<PT>(s);
#pragma unroll
    for (int t = 0; t < XPL; t++) {
        int n = lane + WAVE * t;
        if (n < n1) s[lat<PT>(n)] = v[t];
    }
    if constexpr (MODE == X_TRIAL) {
        // get_trial_state! (timestepping.jl:111-118): rhs = A Γ; rhs += (β1 dt) N1;
        // rhs += (β2 dt) N0; rhs += rhsbc, with rhsbc = apply_bc!(0, get_bc!(Γ[:, lev+1]), fac),
        // then the first x-transform of Ainv.  A = I - qc C'C with qc = dt/(2 Re hc^2)
        // (laplacian.jl:73-81: the eigenvalues 1 - qc Λ of get_A are those of
        // this 5-point stencil with zero Dirichlet values), so A Γ is formed
        // from the rows j-1, j, j+1 of Γ itself.
        const int e = a.lev0 + b;
        const int lev = en.lev;
        const bool hasbc = lev < en.L - 1;
        const double* gc = a.gam + (long)(e + 1) * a.lstride;
        const double hc = level_hc(a.g.h, lev);
        double dt = a.dt, Re = a.Re, c1 = a.c1, c2 = a.c2;
        if (a.ipar) {
            const InstPar& ip = a.ipar[en.inst];
            dt = ip.dt; Re = ip.Re; c1 = ip.c1; c2 = ip.c2;
        }
        const double qc = dt / ((2.0 * Re) * (hc * hc));
        const double fac = 0.25 * dt / (Re * (hc * hc));  // timestepping.jl:104
        const bool rb = (j == 1), rt = (j == a.g.ny - 1);
        double cl = 0.0, cr = 0.0;
        if (hasbc) {
            cl = fac * bc_col(gc, a.g, a.g.nx / 4, 1 + j);
            cr = fac * bc_col(gc, a.g, 3 * a.g.nx / 4, 1 + j);
        }
        const bool rowbc = hasbc && (rb || rt);
        wave_sync();
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            const int n = lane + WAVE * t;
            if (n < n1) {
                const double gw = n > 0 ? s[lat<PT>(n - 1)] : 0.0;
                const double ge = n < n1 - 1 ? s[lat<PT>(n + 1)] : 0.0;
                // A Γ = Γ - qc (4 Γ - neighbours)
                double rhs = v[t] - qc * (4.0 * v[t] - ((gw + ge) + (vm[t] + vp[t])));
                rhs += c1 * n1v[t];
                rhs += c2 * u[t];
                if (hasbc) {
                    // rhsbc = apply_bc!(0, bc, fac): row terms, then column terms
                    const int i = n + 1;
                    if (rowbc || i == 1 || i == a.g.nx - 1) {
                        double rbc = 0.0;
                        if (rowbc) {
                            if (rb) rbc += fac * bc_row(gc, a.g, a.g.ny / 4, 1 + i);
                            if (rt) rbc += fac * bc_row(gc, a.g, 3 * a.g.ny / 4, 1 + i);
                        }
                        if (i == 1) rbc += cl;
                        if (i == a.g.nx - 1) rbc += cr;
                        rhs += rbc;
                    }
                }
                v[t] = rhs;
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            const int n = lane + WAVE * t;
            if (n < n1) s[lat<PT>(n)] = v[t];
        }
    }
    wave_sync();
    dst_line<PT>(s, a.pl, tb, lane);
    double* dst = a.dst + b * a.lstride + (long)line * n1;
    if constexpr (MODE == X_SUB) {
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            int n = lane + WAVE * t;
            if (n < n1) v[t] = dst[n];
        }
#pragma unroll
        for (int t = 0; t < XPL; t++) {
            int n = lane + WAVE * t;
            if (n < n1) dst[n] = v[t] - s[lat<PT>(n)];
        }
    } else {
        if (!a.keep || line == a.keep0 || line == a.keep1) {
#pragma unroll
            for (int t = 0; t < XPL; t++) {
                int n = lane + WAVE * t;
                if (n < n1) dst[n] = s[lat<PT>(n)];
            }
        }
        if (MODE == X_PLAIN && a.edge && lane < 2) {
            // the two columns get_bc! samples for the next finer level, beside each other
            const int ie = (lane == 0 ? a.g.nx / 4 : 3 * a.g.nx / 4) - 1;
            a.edge[b * a.edstride + 2 * line + lane] = s[lat<PT>(ie)];
        }
    }
}

// dynamic LDS above 64 KiB needs the attribute once per kernel instance
template <typename K>
static void allow_big_lds(K kern)
{
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <int MODE, int PT, int XW>
static void launch_x_t(const XArgs& a, hipStream_t s)
{
    static bool once = (allow_big_lds(k_xpass<MODE, PT, XW>), true);
    (void)once;
    const int nlines = a.lsel ? a.lhi - a.llo + 1 : a.g.n2;
    dim3 grid((nlines + XW - 1) / XW, a.nlev);
    size_t lds = sizeof(double) * ((size_t)tab_doubles_t<PT>(a.pl) + (size_t)XW * slot_doubles_t<PT>(a.pl));
    hipLaunchKernelGGL((k_xpass<MODE, PT, XW>), grid, dim3(XW * WAVE), lds, s, a);
}

// plain x-passes over a batch may run in larger work-groups (one copy of the
// tables in LDS for more lines: more waves per CU); the others use XWAVES
template <int MODE, int PT>
static void launch_x_w(const XArgs& a, hipStream_t s)
{
    if constexpr (MODE == X_PLAIN) {
        if (a.nlev > 1) {
            switch (a.waves) {
            case 8: launch_x_t<MODE, PT, 8>(a, s); return;
            case 12: launch_x_t<MODE, PT, 12>(a, s); return;
            case 14: launch_x_t<MODE, PT, 14>(a, s); return;
            default: break;
            }
        }
    }
    launch_x_t<MODE, PT, XWAVES>(a, s);
}

template <int MODE>
static void launch_x_m(const XArgs& a, hipStream_t s)
{
    const int P = a.pl.kind == 0 ? a.pl.P : 0;
    switch (P) {
    case 512: launch_x_w<MODE, 512>(a, s); break;
    case 400: launch_x_w<MODE, 400>(a, s); break;
    case 300: launch_x_w<MODE, 300>(a, s); break;
    case 256: launch_x_w<MODE, 256>(a, s); break;
    case 200: launch_x_w<MODE, 200>(a, s); break;
    case 150: launch_x_w<MODE, 150>(a, s); break;
    case 128: launch_x_w<MODE, 128>(a, s); break;
    case 100: launch_x_w<MODE, 100>(a, s); break;
    default: launch_x_w<MODE, 0>(a, s); break;
    }
}

size_t xpass_lds_bytes(const DstPlan& pl, int waves)
{
    return (pl.kind == 0 && pl.P == 512)
               ? sizeof(double) * ((size_t)tab_doubles_t<512>(pl) + (size_t)waves * slot_doubles_t<512>(pl))
               : sizeof(double) * ((size_t)tab_doubles_t<0>(pl) + (size_t)waves * slot_doubles_t<0>(pl));
}

void launch_xpass(int mode, const XArgs& a, hipStream_t s)
{
    switch (mode) {
    case X_PLAIN: launch_x_m<X_PLAIN>(a, s); break;
    case X_ROT: launch_x_m<X_ROT>(a, s); break;
    case X_SUB: launch_x_m<X_SUB>(a, s); break;
    default: launch_x_m<X_TRIAL>(a, s); break;
    }
}

// ===========================================================================
// DST y-pass, fused: forward DST along y, eigenvalue scaling, DST along y.
// A workgroup stages TI adjacent columns (all rows) of one level in LDS.
// ===========================================================================
constexpr int YWAVES = 4;

// 1/d for the eigenvalue scaling: v_rcp_f64 refined by two Newton steps
// (each squares the relative error), within an ulp or two of a division
__device__ __forceinline__ double recip(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// eigenvalue scaling of dst_inv! (dst-inversion.jl:46-48): w ./= Λ̃; w .*= scale
// with Λ = -2 (cos(pi i/nx) + cos(pi j/ny) - 2) (laplacian.jl:17) and Λ̃ per
// operator (laplacian.jl:69, 74-75).  Applied as one product by the factor
// scale / Λ̃: Δinv 1/Λ, A (1 - Λ dt/(2 Re hc^2)) (the reference divides by its
// reciprocal), Ainv 1/(1 + Λ dt/(2 Re hc^2)); qc = dt / (2 Re hc^2).
__device__ __forceinline__ double eig_factor(double cxk, double cyj, int op, double qc, double scale)
{
    const double lam = -2.0 * ((cxk + cyj) - 2.0);
    if (op == Y_A) return (1.0 - lam * qc) * scale;
    const double d = (op == Y_DINV) ? lam : 1.0 + lam * qc;
    return recip(d) * scale;
}

// dst1024_line output of the y-pass: the first transform of a column stores
// its outputs scaled (op >= 0), the second stores them as they are (op < 0).
// cyp holds cos(pi n/ny) at gsl(n) (and 1 at the guard n = 0, whose value 0
// scales to 0).
struct YOut1024 {
    double* S;
    const double* cyp;
    double cxk, qc, scale;
    int op;
    __device__ __forceinline__ void operator()(int L, int i, double ye, double yo) const
    {
        const int q = 18 * L + 2 * i;  // gsl(16 L + 2 i)
        if (op >= 0) {
            const double2 cy = *reinterpret_cast<const double2*>(cyp + q);
            ye *= eig_factor(cxk, cy.x, op, qc, scale);
            yo *= eig_factor(cxk, cy.y, op, qc, scale);
        }
        *reinterpret_cast<double2*>(S + q) = make_double2(ye, yo);
    }
};

// doubles of the LDS cos(pi j/ny) table
template <int PT>
__host__ __device__ inline int cy_doubles_t(int n2)
{
    if constexpr (PT == 512) return f1024::gsl(f1024::M) + 2;
    else return (n2 + 1) & ~1;
}

// LDS slot (double2 units) of the staged boundary pair e: one pad per 8, so
// that lanes reading pairs 8 apart (consecutive-row threads) hit distinct banks
__host__ __device__ inline int ep_slot(int e) { return e + (e >> 3); }

// dst1024_split output of the y-pass: as YOut1024, one value at a time
// (output n = 16 L + nl of the wavefront's parity)
struct YOutSplit {
    double* S;
    const double* cyp;
    double cxk, qc, scale;
    int op;
    __device__ __forceinline__ void operator()(int L, int nl, double y) const
    {
        const int q = 18 * L + nl;  // gsl(16 L + nl)
        if (op >= 0) y *= eig_factor(cxk, cyp[q], op, qc, scale);
        S[q] = y;
    }
};

// LDS doubles of one column of the split y-pass: the line, the slot the
// first transform writes (the second writes the line again), and the complex
// work arrays of the column's two wavefronts
constexpr int YSPLIT_COL = 2 * f1024::SLOT + 2 * fsplit::ZS;

// BC: the launch carries boundary terms (a.bcr != null); a separate instance
// so that launches without them keep their register budget.  SPLIT (1023-point
// lines, YW = 2 TI): each column is transformed by two wavefronts at once
// (dst1024_split), which puts two wavefronts on every SIMD of a single-level
// launch instead of one.
// YSRC: the launch reads some rows from a.ysrc (a separate instance, so that
// launches without keep their register budget and occupancy).
template <int TI, int PT, int YW, bool BC, bool SPLIT = false, bool YSRC = false>
__global__ void __launch_bounds__(YW * WAVE) k_ypass(YArgs a)
{
    static_assert(!SPLIT || (PT == 512 && YW == 2 * TI), "the split y-pass takes two wavefronts per 1023-point column");
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w = tid / WAVE;
    const int n1 = a.g.n1, n2 = line_len<PT>(a.g.n2);
    // logical block = (level, column tile), consecutive tiles on one XCD
    const int ntiles = (n1 + TI - 1) / TI;
    const int lin = xcd_block(blockIdx.x, gridDim.x / NXCD);
    if (lin >= ntiles * a.nlev) return;
    const int b = lin / ntiles;
    const Entry en = batch_entry(b, a.lev0, a.levmod, a.fixlev, 1 << 30, a.ipar);
    if (!en.active) return;  // the whole work-group: a level its instance does not have
    const int i0 = (lin - b * ntiles) * TI;
    const int ld = SPLIT ? YSPLIT_COL : slot_doubles_t<PT>(a.pl);
    double* base = a.buf + b * a.lstride;
    // cos(pi i/nx) of this wave's columns, loaded with the tile (not after it)
    constexpr int NCOL = (TI + YW - 1) / YW;
    double cxw[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; k++) {
        const int c = w + YW * k;
        cxw[k] = (c < TI && i0 + c < n1) ? a.cx[i0 + c] : 0.0;
    }
    // tile load: thread (r0, c) moves rows r0, r0 + RPI, ... of column c; CH
    // loads per thread are in flight before their LDS writes.  The BC
    // instance gives each thread groups of G consecutive rows instead
    // (G r0 + G RPI g ..), whose column terms need G/2 + 1 staged boundary
    // pairs per group; a wave still covers 16 rows per load either way.  G is
    // CH for the guarded 1023-point layout, whose padding keeps lanes CH rows
    // apart on distinct LDS banks, and 2 for the plain layouts.
    constexpr int RPI = YW * WAVE / TI;
    constexpr int CH = SPLIT ? 8 : 16;
    constexpr int G = PT == 512 ? CH : 2;
    static_assert(CH % G == 0 && G % 2 == 0, "the column terms take rows in pairs");
    const int c0 = tid % TI, r0 = tid / TI;
    const int ic = i0 + c0;
    const bool colok = ic < n1;
    // row of load t of the chunk that starts at row rb (a multiple of RPI CH)
    auto trow = [&](int rb, int t) -> int {
        if constexpr (BC) return rb + G * r0 + G * RPI * (t / G) + (t % G);
        else return rb + r0 + RPI * t;
    };
    // LDS: [DST tables][cos(pi j/ny) table][TI column slots]
    const int ntab = SPLIT ? fsplit::TAB : tab_doubles_t<PT>(a.pl);
    // the 1023-point tables and the cos table are loaded in two phases around
    // the boundary and tile loads below (all in flight together)
    constexpr int NT = YW * WAVE;
    constexpr int CYN = (f1024::M - 1 + NT - 1) / NT;
    std::conditional_t<SPLIT, SplitTabRegs<NT>, Tab1024Regs<NT>> TR;
    double cyr[CYN];
    DstTab tb{};
    SplitTab tbs{};
    double* cyl = lds + ntab;
    if constexpr (SPLIT) {
        splittab_issue(TR, a.pl, tid);
#pragma unroll
        for (int q = 0; q < CYN; q++) cyr[q] = a.cy[min(tid + NT * q, n2 - 1)];
    } else if constexpr (PT == 512) {
        tab1024_issue(TR, a.pl, tid);
#pragma unroll
        for (int q = 0; q < CYN; q++) cyr[q] = a.cy[min(tid + NT * q, n2 - 1)];
    } else {
        tb = load_tab_t<PT, NT>(lds, a.pl, tid);
        for (int k = tid; k < n2; k += NT) cyl[k] = a.cy[k];
    }
    double* tile = cyl + cy_doubles_t<PT>(n2);
    if (tid < TI) {
        set_guards<PT>(tile + tid * ld);
        if (SPLIT) set_guards<PT>(tile + tid * ld + f1024::SLOT);
    }
    // apply_bc! (mg-utils.jl:43-60) with get_bc! (mg-utils.jl:7-32) of the next
    // coarser level, carried through the x-transform.  With S(t) = sin(pi t/nx),
    // C(t) = cos(pi t/nx), line k of the x-transformed field gets
    //   2 fac S(k) (bcL(j) + (-1)^(k+1) bcR(j))          at every row j (columns i = 1, nx-1),
    //   2 fac sum_i S(i k) bcB(i)  at j = 1,  2 fac sum_i S(i k) bcT(i)  at j = ny-1,
    // where, for the coarse row c_0..c_{nx/2} that get_bc! interpolates,
    //   sum_i S(i k) bc(i) = (1 + C(k)) sum_{m=1}^{nx/2-1} S(2 m k) c_m + S(k) (c_0 + (-1)^(k+1) c_{nx/2}) / 2
    // (the odd samples are neighbour means and S(a - b) + S(a + b) = 2 S(a) C(b)).
    const int nx = a.g.nx;
    const bool bc = BC && en.lev < en.L - 1;
    const int ncm = (nx / 2 - 1 + WAVE - 1) / WAVE;  // iterations of the row sums this grid needs
    // LDS after the tile slots: the coarse (L, R) column pairs of lines
    // ny/4 - 1 .. 3ny/4 - 1, pair e at ep_slot(e)
    double2* edl = reinterpret_cast<double2*>(tile + TI * ld);
    const int nep = a.g.ny / 2 + 1;
    double* rowl = reinterpret_cast<double*>(edl + ep_slot(a.g.ny / 2) + 1);  // split: (addB, addT) per column
    constexpr int EPT = (MAXM / 2 + 1 + YW * WAVE - 1) / (YW * WAVE);  // pairs staged per thread
    constexpr int NCM = (MAXM / 2 + WAVE - 1) / WAVE;                  // coarse-row samples per lane
    double2 ep[EPT];
    double addB[NCOL], addT[NCOL];  // row terms of this wave's columns
    double cB[NCM], cT[NCM], cB0 = 0.0, cT0 = 0.0, cBn = 0.0, cTn = 0.0, Sk[NCOL];
    // per column: sin and cos of the lane's first angle 2 pi m k / nx (m = lane + 1)
    // and of the step 2 pi 64 k / nx, as loaded; bit r of sgn[kk]: entry r is negated
    double rsv[NCOL][4];
    unsigned sgn[NCOL];
    double ek = 0.0, esg = 0.0;
    // (1) every load of the boundary data, issued before the tile's so that
    // the row sums wait for these alone while the tile is in flight
    if (bc) {
        const double2* E2 = reinterpret_cast<const double2*>(a.edge + b * a.edstride) + (a.g.ny / 4 - 1);
#pragma unroll
        for (int q = 0; q < EPT; q++) {
            const int e = tid + YW * WAVE * q;
            ep[q] = e < nep ? E2[e] : make_double2(0.0, 0.0);
        }
        const double* rB = a.bcr + b * a.bcstride + (long)n1 * (a.g.ny / 4 - 1) + (nx / 4 - 1);
        const double* rT = a.bcr + b * a.bcstride + (long)n1 * (3 * a.g.ny / 4 - 1) + (nx / 4 - 1);
#pragma unroll
        for (int q = 0; q < NCM; q++) {
            const int m = lane + 1 + WAVE * q;
            cB[q] = cT[q] = 0.0;
            if (q < ncm) {
                cB[q] = m < nx / 2 ? rB[m] : 0.0;
                cT[q] = m < nx / 2 ? rT[m] : 0.0;
            }
        }
        cB0 = rB[0]; cT0 = rT[0]; cBn = rB[nx / 2]; cTn = rT[nx / 2];
        if (colok) {
            const int k = ic + 1;
            ek = a.sinx[min(k, nx - k)];  // scaled by 2 fac once loaded (below)
            esg = (k & 1) ? 1.0 : -1.0;
        }
        // S(2 m k) for the row sums of line k = i0 + c + 1, c = w + YW kk,
        // m = lane + 1 + WAVE q: the angles advance by a fixed step, so two
        // table values per lane and two per column start a rotation.  With
        // sin(pi t / nx) = sign * sinx[min(t', nx - t')], t' = t mod nx, sign
        // negative for t mod 2nx >= nx; cos(pi t / nx) = sin(pi (t + nx/2) / nx).
#pragma unroll
        for (int kk = 0; kk < NCOL; kk++) {
            const int k = min(i0 + w + YW * kk, n1 - 1) + 1;
            Sk[kk] = a.sinx[min(k, nx - k)];
            const int u0 = ((lane + 1) * k) % nx, ud = (WAVE * k) % nx;
            const int tv[4] = {2 * u0, 2 * u0 + nx / 2, 2 * ud, 2 * ud + nx / 2};
            sgn[kk] = 0u;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int t = tv[r] >= 2 * nx ? tv[r] - 2 * nx : tv[r];
                const int tt = t >= nx ? t - nx : t;
                rsv[kk][r] = a.sinx[min(tt, nx - tt)];
                sgn[kk] |= (t >= nx ? 1u : 0u) << r;
            }
        }
    }
    // (2) the first chunk of the tile (rows outside rlo..rhi are zero when rsel;
    // rows outside clo..chi come from ysrc when it is given)
    const int rlo = a.rsel ? a.rlo : 0, rhi = a.rsel ? a.rhi : n2 - 1;
    const double* ys = YSRC ? a.ysrc + b * a.lstride : nullptr;
    const double* yb = (YSRC && a.ysub) ? a.ysub + b * a.lstride : nullptr;
    auto load_chunk = [&](int rb, double* vv) {
        if constexpr (!YSRC) {
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int r = trow(rb, t);
                vv[t] = (colok && r >= rlo && r <= rhi) ? base[ic + (long)n1 * r] : 0.0;
            }
        } else {
            // every load of the chunk is issued before the first is used
            bool alt[CH];
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int r = trow(rb, t);
                const bool ok = colok && r >= rlo && r <= rhi;
                alt[t] = ok && (r < a.clo || r > a.chi);
                vv[t] = ok ? (alt[t] ? ys : base)[ic + (long)n1 * r] : 0.0;
            }
            double sb[CH];
#pragma unroll
            for (int t = 0; t < CH; t++) sb[t] = (yb && alt[t]) ? yb[ic + (long)n1 * trow(rb, t)] : 0.0;
#pragma unroll
            for (int t = 0; t < CH; t++)
                if (alt[t]) vv[t] = a.afac * (yb ? vv[t] - sb[t] : vv[t]);
        }
    };
    double v[CH];
    load_chunk(0, v);
    if constexpr (PT == 512) {
        if constexpr (SPLIT) tbs = splittab_store(lds, TR, tid);
        else tb = tab1024_store(lds, TR, tid);
#pragma unroll
        for (int q = 0; q < CYN; q++) {
            const int k = tid + NT * q;
            if (k < n2) cyl[f1024::gsl(k + 1)] = cyr[q];
        }
        if (tid == 0) cyl[f1024::gsl(0)] = 1.0;
    }
    // (3) row sums and the staged column pairs
    if (bc) {
        const double f2 = 2.0 * a.bcfac;
#pragma unroll
        for (int kk = 0; kk < NCOL; kk++) {
            const int k = min(i0 + w + YW * kk, n1 - 1) + 1;
            double v4[4];
#pragma unroll
            for (int r = 0; r < 4; r++) v4[r] = ((sgn[kk] >> r) & 1u) ? -rsv[kk][r] : rsv[kk][r];
            double sn = v4[0], cs = v4[1];
            const double sd = v4[2], cd = v4[3];
            double sB = 0.0, sT = 0.0;
#pragma unroll
            for (int q = 0; q < NCM; q++) {
                sB += sn * cB[q];
                sT += sn * cT[q];
                const double s2 = sn * cd + cs * sd;
                cs = cs * cd - sn * sd;
                sn = s2;
            }
            sB = wave_last(wave_incl_scan(sB, lane));
            sT = wave_last(wave_incl_scan(sT, lane));
            const double sgk = (k & 1) ? 1.0 : -1.0;
            addB[kk] = f2 * ((1.0 + cxw[kk]) * sB + 0.5 * Sk[kk] * (cB0 + sgk * cBn));
            addT[kk] = f2 * ((1.0 + cxw[kk]) * sT + 0.5 * Sk[kk] * (cT0 + sgk * cTn));
        }
        ek *= 2.0 * a.bcfac;
#pragma unroll
        for (int q = 0; q < EPT; q++) {
            const int e = tid + YW * WAVE * q;
            if (e < nep) edl[ep_slot(e)] = ep[q];
        }
        // split: the row terms of column w go with the tile into the column
        // (its two wavefronts both read rows 0 and n2 - 1)
        if (SPLIT && w < TI && lane == 0) {
            rowl[2 * w] = addB[0];
            rowl[2 * w + 1] = addT[0];
        }
        __syncthreads();
    }
    // (4) the tile into LDS, with the column terms
    for (int rb = 0; rb < n2; rb += RPI * CH) {
        if (rb > 0) load_chunk(rb, v);
        if (bc) {
            // get_bc! sample idx = j + 1 = r + 2 of the left and right columns:
            // the even row re = trow(rb, 2u) (idx = 2e, e = re/2 + 1) takes the
            // mean of the staged pairs e and e - 1, the odd row re + 1
            // (idx = 2e + 1) pair e.  A group's rows need pairs e0 - 1 ..
            // e0 - 1 + G/2; those past the last pair (rows beyond the line)
            // are read clamped and not used.
#pragma unroll
            for (int gi = 0; gi < CH / G; gi++) {
                const int e0 = trow(rb, G * gi) / 2 + 1;
                double2 pr[G / 2 + 1];
#pragma unroll
                for (int u = 0; u <= G / 2; u++) pr[u] = edl[ep_slot(min(e0 - 1 + u, nep - 1))];
#pragma unroll
                for (int u = 0; u < G / 2; u++) {
                    const int t = G * gi + 2 * u;
                    const int re = trow(rb, t);
                    if (re < n2) {
                        const double2 ea = pr[u + 1], eb = pr[u];
                        const double bl = 0.5 * (ea.x + eb.x), br = 0.5 * (ea.y + eb.y);
                        v[t] += ek * (bl + esg * br);
                        if (re + 1 < n2) v[t + 1] += ek * (ea.x + esg * ea.y);
                    }
                }
            }
        }
        if (SPLIT && bc) {
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int r = trow(rb, t);
                if (r == 0) v[t] += rowl[2 * c0];
                if (r == n2 - 1) v[t] += rowl[2 * c0 + 1];
            }
        }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            const int r = trow(rb, t);
            if (r < n2) tile[c0 * ld + lat<PT>(r)] = v[t];
        }
    }
    __syncthreads();
    const int lev = en.lev;
    const double hcl = level_hc(a.g.h, lev);
    double dt = a.dt, Re = a.Re;
    if (a.ipar) {
        dt = a.ipar[en.inst].dt;
        Re = a.ipar[en.inst].Re;
    }
    const double qc = dt / ((2.0 * Re) * (hcl * hcl));  // dt / (2 Re h^2), laplacian.jl:69,74
    if constexpr (SPLIT) {
        // column w / 2, outputs of parity w % 2: the first transform reads the
        // line and writes the scaled outputs into the second slot, the second
        // transform reads that slot and writes the line
        const int c = w >> 1, half = w & 1;
        const bool on = i0 + c < n1;
        const double cxk = on ? a.cx[i0 + c] : 0.0;
        double* s0 = tile + c * ld;
        double* s1 = s0 + f1024::SLOT;
        double2* zw = reinterpret_cast<double2*>(s0 + 2 * f1024::SLOT + half * fsplit::ZS);
        if (on) dst1024_split(s0, zw, tbs, lane, half, YOutSplit{s1, cyl, cxk, qc, a.scale, a.op});
        __syncthreads();
        if (on) dst1024_split(s1, zw, tbs, lane, half, YOutSplit{s0, cyl, cxk, qc, a.scale, -1});
    } else {
    // (column, transform) pairs: transform 0 is followed by the eigenvalue
    // scaling; one call site keeps a single DST body in the kernel
#pragma unroll 1
    for (int it = 0; it < 2 * NCOL; it++) {
        const int c = w + YW * (it >> 1);
        if (c >= TI || i0 + c >= n1) continue;
        double* s = tile + c * ld;
        double cxk = cxw[0];
#pragma unroll
        for (int k = 1; k < NCOL; k++)
            if ((it >> 1) == k) cxk = cxw[k];
        if (bc && (it & 1) == 0) {
            double ab = addB[0], at = addT[0];
#pragma unroll
            for (int k = 1; k < NCOL; k++)
                if ((it >> 1) == k) {
                    ab = addB[k];
                    at = addT[k];
                }
            if (lane == 0) {
                s[lat<PT>(0)] += ab;
                s[lat<PT>(n2 - 1)] += at;
            }
            wave_sync();
        }
        if constexpr (PT == 512) {
            dst1024_line(s, tb, lane, YOut1024{s, cyl, cxk, qc, a.scale, (it & 1) == 0 ? a.op : -1});
        } else {
            dst_line<PT>(s, a.pl, tb, lane);
            if ((it & 1) == 0) {
                for (int kj = lane; kj < n2; kj += WAVE) s[kj] *= eig_factor(cxk, cyl[kj], a.op, qc, a.scale);
                wave_sync();
            }
        }
    }
    }
    __syncthreads();
    if (colok) {
        const int wlo = a.wsel ? a.wlo : 0, whi = a.wsel ? a.whi : n2 - 1;
        for (int rb = 0; rb < n2; rb += RPI * CH) {
            double v[CH];
#pragma unroll
            for (int t = 0; t < CH; t++) {
                int r = rb + r0 + RPI * t;
                if (r < n2) v[t] = tile[c0 * ld + lat<PT>(r)];
            }
#pragma unroll
            for (int t = 0; t < CH; t++) {
                int r = rb + r0 + RPI * t;
                if (r >= wlo && r <= whi) base[ic + (long)n1 * r] = v[t];
            }
        }
    }
}

template <int PT>
static size_t ypass_lds_bytes_t(const DstPlan& pl, int n2, int tile)
{
    return sizeof(double) * ((size_t)tab_doubles_t<PT>(pl) + (size_t)cy_doubles_t<PT>(n2) +
                             (size_t)tile * slot_doubles_t<PT>(pl));
}

size_t ypass_lds_bytes(const DstPlan& pl, int n2, int tile)
{
    return (pl.kind == 0 && pl.P == 512) ? ypass_lds_bytes_t<512>(pl, n2, tile) : ypass_lds_bytes_t<0>(pl, n2, tile);
}

// the split y-pass: 4-column tiles of 1023-point columns, 8 wavefronts
static void launch_y_split(const YArgs& a, hipStream_t s)
{
    static bool once = (allow_big_lds(k_ypass<4, 512, 8, false, true>), allow_big_lds(k_ypass<4, 512, 8, true, true>),
                        allow_big_lds(k_ypass<4, 512, 8, false, true, true>),
                        allow_big_lds(k_ypass<4, 512, 8, true, true, true>), true);
    (void)once;
    const int nblocks = ((a.g.n1 + 3) / 4) * a.nlev;
    dim3 grid(NXCD * xcd_nper(nblocks));
    const size_t lds = ypass_split_lds_bytes(a.g.ny, a.bcr != nullptr);
    if (a.ysrc) {
        if (a.bcr) hipLaunchKernelGGL((k_ypass<4, 512, 8, true, true, true>), grid, dim3(8 * WAVE), lds, s, a);
        else hipLaunchKernelGGL((k_ypass<4, 512, 8, false, true, true>), grid, dim3(8 * WAVE), lds, s, a);
    } else {
        if (a.bcr) hipLaunchKernelGGL((k_ypass<4, 512, 8, true, true>), grid, dim3(8 * WAVE), lds, s, a);
        else hipLaunchKernelGGL((k_ypass<4, 512, 8, false, true>), grid, dim3(8 * WAVE), lds, s, a);
    }
}

size_t ypass_split_lds_bytes(int ny, bool bc)
{
    return sizeof(double) * ((size_t)fsplit::TAB + (size_t)cy_doubles_t<512>(ny - 1) + 4 * (size_t)YSPLIT_COL) +
           (bc ? sizeof(double2) * (size_t)(ep_slot(ny / 2) + 1) + 8 * sizeof(double) : 0);
}

template <int TI, int PT, int YW>
static void launch_y_t(const YArgs& a, hipStream_t s)
{
    static bool once = (allow_big_lds(k_ypass<TI, PT, YW, false>), allow_big_lds(k_ypass<TI, PT, YW, true>),
                        allow_big_lds(k_ypass<TI, PT, YW, false, false, true>),
                        allow_big_lds(k_ypass<TI, PT, YW, true, false, true>), true);
    (void)once;
    const int nblocks = ((a.g.n1 + TI - 1) / TI) * a.nlev;
    dim3 grid(NXCD * xcd_nper(nblocks));
    const size_t lds = ypass_lds_bytes_t<PT>(a.pl, a.g.n2, TI) +
                       (a.bcr ? sizeof(double2) * (size_t)(ep_slot(a.g.ny / 2) + 1) : 0);
    if (a.ysrc) {
        if (a.bcr) hipLaunchKernelGGL((k_ypass<TI, PT, YW, true, false, true>), grid, dim3(YW * WAVE), lds, s, a);
        else hipLaunchKernelGGL((k_ypass<TI, PT, YW, false, false, true>), grid, dim3(YW * WAVE), lds, s, a);
    } else {
        if (a.bcr) hipLaunchKernelGGL((k_ypass<TI, PT, YW, true>), grid, dim3(YW * WAVE), lds, s, a);
        else hipLaunchKernelGGL((k_ypass<TI, PT, YW, false>), grid, dim3(YW * WAVE), lds, s, a);
    }
}

template <int TI, int YW>
static void launch_y_m(const YArgs& a, hipStream_t s)
{
    const int P = a.pl.kind == 0 ? a.pl.P : 0;
    switch (P) {
    case 512: launch_y_t<TI, 512, YW>(a, s); break;
    case 400: launch_y_t<TI, 400, YW>(a, s); break;
    case 300: launch_y_t<TI, 300, YW>(a, s); break;
    case 256: launch_y_t<TI, 256, YW>(a, s); break;
    case 200: launch_y_t<TI, 200, YW>(a, s); break;
    case 150: launch_y_t<TI, 150, YW>(a, s); break;
    case 128: launch_y_t<TI, 128, YW>(a, s); break;
    case 100: launch_y_t<TI, 100, YW>(a, s); break;
    default: launch_y_t<TI, 0, YW>(a, s); break;
    }
}


// ---------------------------------------------------------------------------
// The fused y-pass (no boundary terms) over many column tiles, as persistent
// work-groups: a work-group loads the tables once and walks a list of tiles,
// the global loads of its next tile in flight while the current tile is
// transformed, and the stores of a tile draining while the next one is.
// The arithmetic of a tile is k_ypass's (4 columns, a wavefront each).
// Work-group bid runs on XCD bid % NXCD; the XCDs take contiguous ranges of
// tiles and a work-group every wpx-th tile of its XCD's range, so that the
// work-groups of one XCD work on neighbouring tiles at the same time.
// ---------------------------------------------------------------------------
template <int PT>
__global__ void __launch_bounds__(4 * WAVE) k_ypass_pipe(YArgs a)
{
    constexpr int TI = 4, YW = 4, NT = YW * WAVE, RPI = NT / TI, CH = 16;
    static_assert(RPI * CH >= MAXM - 1, "one chunk of loads covers a column");
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w = tid / WAVE;
    const int n1 = a.g.n1, n2 = line_len<PT>(a.g.n2);
    const int ntiles = (n1 + TI - 1) / TI;
    const int total = ntiles * a.nlev;
    const int wpx = gridDim.x / NXCD;
    const int xcd = blockIdx.x % NXCD, li = blockIdx.x / NXCD;
    const int tpx = (total + NXCD - 1) / NXCD;
    const int t_end = min(total, (xcd + 1) * tpx);
    int t = xcd * tpx + li;
    if (t >= t_end) return;  // the whole work-group, before any barrier
    const int ld = slot_doubles_t<PT>(a.pl);
    const int ntab = tab_doubles_t<PT>(a.pl);
    const int c0 = tid % TI, r0 = tid / TI;
    // tables (once per work-group)
    constexpr int CYN = (f1024::M - 1 + NT - 1) / NT;
    DstTab tb{};
    double* cyl = lds + ntab;
    if constexpr (PT == 512) {
        Tab1024Regs<NT> TR;
        double cyr[CYN];
        tab1024_issue(TR, a.pl, tid);
#pragma unroll
        for (int q = 0; q < CYN; q++) cyr[q] = a.cy[min(tid + NT * q, n2 - 1)];
        tb = tab1024_store(lds, TR, tid);
#pragma unroll
        for (int q = 0; q < CYN; q++) {
            const int k = tid + NT * q;
            if (k < n2) cyl[f1024::gsl(k + 1)] = cyr[q];
        }
        if (tid == 0) cyl[f1024::gsl(0)] = 1.0;
    } else {
        tb = load_tab_t<PT, NT>(lds, a.pl, tid);
        for (int k = tid; k < n2; k += NT) cyl[k] = a.cy[k];
    }
    double* tile = cyl + cy_doubles_t<PT>(n2);
    if (tid < TI) set_guards<PT>(tile + tid * ld);
    // tile t: entry b = t / ntiles, columns i0 .. i0 + TI - 1
    auto active = [&](int tt) -> bool {
        const int b = tt / ntiles;
        return batch_entry(b, a.lev0, a.levmod, a.fixlev, 1 << 30, a.ipar).active;
    };
    auto load = [&](int tt, double* v) {
        const int b = tt / ntiles, ic = (tt - b * ntiles) * TI + c0;
        const double* base = a.buf + b * a.lstride;
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int r = r0 + RPI * k;
            v[k] = (ic < n1 && r < n2) ? base[ic + (long)n1 * r] : 0.0;
        }
    };
    double v[CH];
    bool act = active(t);
    if (act) load(t, v);
    for (;;) {
        // the current tile into LDS
        if (act) {
#pragma unroll
            for (int k = 0; k < CH; k++) {
                const int r = r0 + RPI * k;
                if (r < n2) tile[c0 * ld + lat<PT>(r)] = v[k];
            }
        }
        __syncthreads();
        // the next tile's loads in flight during this tile's transforms
        const int tn = t + wpx;
        const bool has_next = tn < t_end;
        const bool actn = has_next && active(tn);
        if (actn) load(tn, v);
        if (act) {
            const int b = t / ntiles, i0 = (t - b * ntiles) * TI;
            const Entry en = batch_entry(b, a.lev0, a.levmod, a.fixlev, 1 << 30, a.ipar);
            const double hcl = level_hc(a.g.h, en.lev);
            double dt = a.dt, Re = a.Re;
            if (a.ipar) {
                dt = a.ipar[en.inst].dt;
                Re = a.ipar[en.inst].Re;
            }
            const double qc = dt / ((2.0 * Re) * (hcl * hcl));  // dt / (2 Re h^2), laplacian.jl:69,74
            const int c = w;  // TI == YW: one column per wavefront
            if (i0 + c < n1) {
                const double cxk = a.cx[i0 + c];
                double* s = tile + c * ld;
#pragma unroll 1
                for (int it = 0; it < 2; it++) {
                    if constexpr (PT == 512) {
                        dst1024_line(s, tb, lane, YOut1024{s, cyl, cxk, qc, a.scale, it == 0 ? a.op : -1});
                    } else {
                        dst_line<PT>(s, a.pl, tb, lane);
                        if (it == 0) {
                            for (int kj = lane; kj < n2; kj += WAVE) s[kj] *= eig_factor(cxk, cyl[kj], a.op, qc, a.scale);
                            wave_sync();
                        }
                    }
                }
            }
        }
        __syncthreads();
        // this tile's stores (they drain during the next tile's transforms)
        if (act) {
            const int b = t / ntiles, ic = (t - b * ntiles) * TI + c0;
            double* base = a.buf + b * a.lstride;
            if (ic < n1) {
                double o[CH];
#pragma unroll
                for (int k = 0; k < CH; k++) {
                    const int r = r0 + RPI * k;
                    if (r < n2) o[k] = tile[c0 * ld + lat<PT>(r)];
                }
#pragma unroll
                for (int k = 0; k < CH; k++) {
                    const int r = r0 + RPI * k;
                    if (r < n2) base[ic + (long)n1 * r] = o[k];
                }
            }
        }
        if (!has_next) break;
        __syncthreads();  // the slots are read before the next tile overwrites them
        t = tn;
        act = actn;
    }
}

template <int PT>
static void launch_y_pipe_t(const YArgs& a, int nwg, hipStream_t s)
{
    static bool once = (allow_big_lds(k_ypass_pipe<PT>), true);
    (void)once;
    const size_t lds = ypass_lds_bytes_t<PT>(a.pl, a.g.n2, 4);
    hipLaunchKernelGGL((k_ypass_pipe<PT>), dim3(nwg), dim3(4 * WAVE), lds, s, a);
}

// tile 4: four waves, a column each; 8: four waves, two columns each; 12:
// twelve waves, a column each (one copy of the tables for three times the
// columns of tile 4)
void launch_ypass(const YArgs& a, int tile, hipStream_t s)
{
    // pipelined persistent work-groups (tile option 4 with pipe): launches of
    // more 4-column tiles than two work-groups per CU hold at once
    const int ntiles4 = ((a.g.n1 + 3) / 4) * a.nlev;
    if (tile == 4 && a.pipe && !a.bcr && !a.ysrc && ntiles4 > 2 * 256) {
        const int nwg = NXCD * 64;  // two per CU of the 256
        const int P = a.pl.kind == 0 ? a.pl.P : 0;
        switch (P) {
        case 512: launch_y_pipe_t<512>(a, nwg, s); return;
        case 400: launch_y_pipe_t<400>(a, nwg, s); return;
        case 300: launch_y_pipe_t<300>(a, nwg, s); return;
        case 256: launch_y_pipe_t<256>(a, nwg, s); return;
        case 200: launch_y_pipe_t<200>(a, nwg, s); return;
        case 150: launch_y_pipe_t<150>(a, nwg, s); return;
        case 128: launch_y_pipe_t<128>(a, nwg, s); return;
        case 100: launch_y_pipe_t<100>(a, nwg, s); return;
        default: launch_y_pipe_t<0>(a, nwg, s); return;
        }
    }
    if (tile == 4 && a.split && a.pl.kind == 0 && a.pl.P == 512 && ntiles4 <= 256 &&
        ypass_split_lds_bytes(a.g.ny, a.bcr != nullptr) <= 160 * 1024) {
        launch_y_split(a, s);
        return;
    }
    if (tile == 4) launch_y_m<4, YWAVES>(a, s);
    else if (tile == 12) launch_y_m<12, 12>(a, s);
    else launch_y_m<8, YWAVES>(a, s);
}

// ===========================================================================
// nonlinear! (nonlinear.jl:168-201), fused: avg_flux (:17-38),
// direct_product_loops! (:73-114), rot! (curl.jl:22-34) and the 1/(4 hc^2)
// scaling.  Boundary circulation is get_bc!(Γ[:, lev+1]) .* 0.25.
// A thread takes a vertical strip of NL_JS output Γ points.
//
// The fluxes come from the state's q, or are formed from its ψ by curl!
// (the values curl_qx / curl_qy give, which are what curl! stores in q): a
// strip whose stencil stays off the edges reads ψ directly, an edge strip
// goes through curl_qx / curl_qy with the boundary values.  The products
// are the same expression in every form.
// ===========================================================================
constexpr int NL_JS = 8;  // output rows per thread (vertical strip)

// Flux and circulation sources of one level.  qx / qy: the flux without the
// base flux; gam: Γ with its boundary ring; gb: ΓB with its ring (linearized).
struct NLLevel {
    const GridDev* g;
    const double* __restrict__ G;
    const double* __restrict__ Gc;  // Γ of the next coarser level (hasbc)
    const double* __restrict__ GBl;
    bool hasbc;
    __device__ __forceinline__ double bcv(int k) const { return hasbc ? get_bc_val(Gc, *g, k) * 0.25 : 0.0; }
    // a circulation field extended by the boundary ring get_bc!(Γ[:, lev+1]) .* 0.25:
    // X(ii, jj), ii = 0..nx, jj = 0..ny (corners are never read)
    __device__ __forceinline__ double ext(const double* __restrict__ F, int ii, int jj) const
    {
        if (jj == 0) return bcv(g->BOT + ii + 1);
        if (jj == g->ny) return bcv(g->TOP + ii + 1);
        if (ii == 0) return bcv(g->LEFT + jj + 1);
        if (ii == g->nx) return bcv(g->RIGHT + jj + 1);
        return F[(ii - 1) + (long)g->n1 * (jj - 1)];
    }
    __device__ __forceinline__ double gam(int ii, int jj) const { return ext(G, ii, jj); }
    __device__ __forceinline__ double gb(int ii, int jj) const { return ext(GBl, ii, jj); }
};
struct NLFromQ : NLLevel {
    const double* __restrict__ q;
    __device__ __forceinline__ double qx(int ii, int jj) const { return q[(ii - 1) + (long)(g->nx + 1) * (jj - 1)]; }
    __device__ __forceinline__ double qy(int ii, int jj) const { return q[g->nu + (ii - 1) + (long)g->nx * (jj - 1)]; }
};
// fluxes from ψ and Γ with its ring, every value by selection rather than
// branches (curl_qx_sel, curl_qy_sel, get_bc_val_sel), so that all loads of
// an edge strip are in flight together; the values are the branching forms'
struct NLFromPsiSel : NLLevel {
    const double* __restrict__ ps;   // ψ of the level
    const double* __restrict__ psc;  // ψ of the next coarser level (hasbc)
    __device__ __forceinline__ double qx(int ii, int jj) const { return curl_qx_sel(ps, psc, hasbc, *g, ii, jj); }
    __device__ __forceinline__ double qy(int ii, int jj) const { return curl_qy_sel(ps, psc, hasbc, *g, ii, jj); }
    __device__ __forceinline__ double gam(int ii, int jj) const
    {
        const GridDev& gg = *g;
        const bool ring = jj == 0 || jj == gg.ny || ii == 0 || ii == gg.nx;
        const int k = jj == 0 ? gg.BOT + ii + 1
                    : jj == gg.ny ? gg.TOP + ii + 1
                    : ii == 0 ? gg.LEFT + jj + 1
                    : ii == gg.nx ? gg.RIGHT + jj + 1 : gg.LEFT + 1;
        const int ic = min(max(ii, 1), gg.nx - 1), jc = min(max(jj, 1), gg.ny - 1);
        const double vin = G[(ic - 1) + (long)gg.n1 * (jc - 1)];
        const double vr = get_bc_val_sel(Gc, G, hasbc, gg, k) * 0.25;
        return ring ? vr : vin;
    }
};
// One strip: output points (i-1, j-1) for j = j0..j1.  UNROLL: the row loop
// unrolled (the common strips) or rolled (edge strips, whose sources carry
// the boundary branches).
template <bool LIN, int UNROLL, class SRC>
__device__ __forceinline__ void nl_strip(const NLArgs& a, const SRC& S, const double* __restrict__ q0, bool u0,
                                         double cx0, double cy0, double hcl, int i, int j0, int j1,
                                         const double* __restrict__ QBl, double* __restrict__ out)
{
    const GridDev& g = *S.g;
    const long nu = g.nu;
    const int nx = g.nx, ny = g.ny, n1 = g.n1;
    // flux components with the base flux: x-flux (ii, jj) and y-flux (ii, jj)
    auto QXV = [&](int ii, int jj, double& f, double& f0) {
        f = S.qx(ii, jj);
        f0 = (LIN || u0) ? cx0 : q0[(ii - 1) + (long)(nx + 1) * (jj - 1)];
    };
    auto QYV = [&](int ii, int jj, double& f, double& f0) {
        f = S.qy(ii, jj);
        f0 = (LIN || u0) ? cy0 : q0[nu + (ii - 1) + (long)nx * (jj - 1)];
    };
    // avg_flux (nonlinear.jl:17-38; linearized :45-65 without the base flux):
    // Qy(ii, jj) over (ii, jj), (ii-1, jj); Qx over (ii, jj), (ii, jj-1)
    auto QY = [&](int ii, int jj) -> double {
        double f, f0, fm, fm0;
        QYV(ii, jj, f, f0);
        QYV(ii - 1, jj, fm, fm0);
        if constexpr (LIN) return f + fm;
        else return ((f + f0) + fm) + fm0;
    };
    // linearized problem: base flux QB (already averaged) and ΓB with the ring
    auto QBX = [&](int ii, int jj) -> double { return QBl[(ii - 1) + (long)(nx + 1) * (jj - 1)]; };
    auto QBY = [&](int ii, int jj) -> double { return QBl[nu + (ii - 1) + (long)nx * (jj - 1)]; };

    // state carried from row j-1 to row j
    double qxr[3], q0r[3];  // x-flux (i-1..i+1, j-1)
    double gr[3], gbr[3];   // Γ(i-2..i, j-1), ΓB(i-2..i, j-1) with the ring
#pragma unroll
    for (int d = 0; d < 3; d++) {
        QXV(i - 1 + d, j0 - 1, qxr[d], q0r[d]);
        gr[d] = S.gam(i - 2 + d, j0 - 1);
        gbr[d] = LIN ? S.gb(i - 2 + d, j0 - 1) : 0.0;
    }
    double qyj = QY(i, j0);  // Qy(i, j)
    double qbyj = LIN ? QBY(i, j0) : 0.0;
    // fq_x(i, j-1): Q x Γ, or Q x ΓB + QB x Γ
    double fqxa;
    if constexpr (LIN)
        fqxa = (qyj * gbr[1] + QY(i, j0 - 1) * S.gb(i - 1, j0 - 2)) + (qbyj * gr[1] + QBY(i, j0 - 1) * S.gam(i - 1, j0 - 2));
    else
        fqxa = qyj * gr[1] + QY(i, j0 - 1) * S.gam(i - 1, j0 - 2);
    const double sl = 1.0 / (4.0 * (hcl * hcl));  // nonlinear.jl:197-198
#pragma unroll UNROLL
    for (int t = 0; t < NL_JS; t++) {
        const int j = j0 + t;
        if (j > j1) break;
        double qxn[3], q0n[3], gn[3], QX[3], gbn[3], qbx[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            QXV(i - 1 + d, j, qxn[d], q0n[d]);
            if constexpr (LIN) {
                QX[d] = qxn[d] + qxr[d];
                qbx[d] = QBX(i - 1 + d, j);
            } else {
                QX[d] = ((qxn[d] + q0n[d]) + qxr[d]) + q0r[d];
            }
        }
        gn[1] = S.gam(i - 1, j);
        if constexpr (LIN) gbn[1] = S.gb(i - 1, j);
        // row ny is the top ring: only X(i-1, ny) is read (the others are
        // evaluated, valid ring entries, and replaced by zero)
        {
            const bool top = j >= ny;
            const double a0 = S.gam(i - 2, j), a2 = S.gam(i, j);
            gn[0] = top ? 0.0 : a0;
            gn[2] = top ? 0.0 : a2;
            if constexpr (LIN) {
                const double b0 = S.gb(i - 2, j), b2 = S.gb(i, j);
                gbn[0] = top ? 0.0 : b0;
                gbn[2] = top ? 0.0 : b2;
            }
        }
        const double qyn = QY(i, j + 1);
        // direct_product_loops!: fqx(i, j) = Qy(i, j+1) Γ(i-1, j) + Qy(i, j) Γ(i-1, j-1);
        // fqy(ii, j) = -(Qx(ii+1, j) Γ(ii, j-1) + Qx(ii, j) Γ(ii-1, j-1)) for ii = i, i-1;
        // the linearized fq is the sum of the (Q, ΓB) and (QB, Γ) products
        double fqxb, fqya, fqyb;
        if constexpr (LIN) {
            const double qbyn = QBY(i, j + 1);
            fqxb = (qyn * gbn[1] + qyj * gbr[1]) + (qbyn * gn[1] + qbyj * gr[1]);
            fqya = ((QX[2] * gbr[2] + QX[1] * gbr[1]) * -1.0) + ((qbx[2] * gr[2] + qbx[1] * gr[1]) * -1.0);
            fqyb = ((QX[1] * gbr[1] + QX[0] * gbr[0]) * -1.0) + ((qbx[1] * gr[1] + qbx[0] * gr[0]) * -1.0);
            qbyj = qbyn;
        } else {
            fqxb = qyn * gn[1] + qyj * gr[1];
            fqya = (QX[2] * gr[2] + QX[1] * gr[1]) * -1.0;
            fqyb = (QX[1] * gr[1] + QX[0] * gr[0]) * -1.0;
        }
        // rot! (curl.jl:22-34) of fq, scaled by 1/(4 hc^2)
        const double r = (fqxa - fqxb) + (fqya - fqyb);
        out[(i - 2) + (long)n1 * (j - 2)] = r * sl;
        fqxa = fqxb;
        qyj = qyn;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            qxr[d] = qxn[d];
            q0r[d] = q0n[d];
            gr[d] = gn[d];
            if constexpr (LIN) gbr[d] = gbn[d];
        }
    }
}

// A strip whose stencil is interior, fluxes from ψ (LIN = false): each flux
// is the interior branch of curl_qx / curl_qy, the ψ rows i-2..i are
// carried from row to row, and the products are nl_strip's.
__device__ __forceinline__ void nl_strip_psi_inner(const double* __restrict__ ps, const double* __restrict__ G,
                                                   const double* __restrict__ q0, bool u0, double cx0, double cy0,
                                                   double hcl, int i, int j0, int j1, const GridDev& g,
                                                   double* __restrict__ out)
{
    const int n1 = g.n1, nx = g.nx;
    const long nu = g.nu;
    auto PS = [&](int ii, int jj) -> double { return ps[(ii - 1) + (long)n1 * (jj - 1)]; };
    auto GA = [&](int ii, int jj) -> double { return G[(ii - 1) + (long)n1 * (jj - 1)]; };
    auto Q0X = [&](int ii, int jj) -> double { return u0 ? cx0 : q0[(ii - 1) + (long)(nx + 1) * (jj - 1)]; };
    auto Q0Y = [&](int ii, int jj) -> double { return u0 ? cy0 : q0[nu + (ii - 1) + (long)nx * (jj - 1)]; };
    // avg_flux's Qy(i, jj) from the ψ row jj - 1 (columns i-2..i): the y-fluxes
    // (i, jj) = ψ(i-1, jj-1) - ψ(i, jj-1) and (i-1, jj) = ψ(i-2, jj-1) - ψ(i-1, jj-1)
    auto QY = [&](const double* pr, int jj) -> double {
        const double f = pr[1] - pr[2], fm = pr[0] - pr[1];
        return ((f + Q0Y(i, jj)) + fm) + Q0Y(i - 1, jj);
    };
    double pm2[3], pm1[3];  // ψ rows j0-2 and j0-1, columns i-2..i
#pragma unroll
    for (int d = 0; d < 3; d++) {
        pm2[d] = PS(i - 2 + d, j0 - 2);
        pm1[d] = PS(i - 2 + d, j0 - 1);
    }
    double qxr[3], q0r[3], gr[3];  // x-flux (i-1..i+1, j-1), Γ(i-2..i, j-1)
#pragma unroll
    for (int d = 0; d < 3; d++) {
        qxr[d] = pm1[d] - pm2[d];
        q0r[d] = Q0X(i - 1 + d, j0 - 1);
        gr[d] = GA(i - 2 + d, j0 - 1);
    }
    double qyj = QY(pm1, j0);
    double fqxa = qyj * gr[1] + QY(pm2, j0 - 1) * GA(i - 1, j0 - 2);
    const double sl = 1.0 / (4.0 * (hcl * hcl));
#pragma unroll
    for (int t = 0; t < NL_JS; t++) {
        const int j = j0 + t;
        if (j > j1) break;
        double pj[3], QX[3], q0n[3], gn[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            pj[d] = PS(i - 2 + d, j);
            q0n[d] = Q0X(i - 1 + d, j);
            gn[d] = GA(i - 2 + d, j);
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double qxn = pj[d] - pm1[d];
            QX[d] = ((qxn + q0n[d]) + qxr[d]) + q0r[d];
            qxr[d] = qxn;
        }
        const double qyn = QY(pj, j + 1);
        const double fqxb = qyn * gn[1] + qyj * gr[1];
        const double fqya = (QX[2] * gr[2] + QX[1] * gr[1]) * -1.0;
        const double fqyb = (QX[1] * gr[1] + QX[0] * gr[0]) * -1.0;
        const double r = (fqxa - fqxb) + (fqya - fqyb);
        out[(i - 2) + (long)n1 * (j - 2)] = r * sl;
        fqxa = fqxb;
        qyj = qyn;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            pm1[d] = pj[d];
            q0r[d] = q0n[d];
            gr[d] = gn[d];
        }
    }
}

// The fluxes from ψ: k_nonlinear_psi takes the output points (i-1, j-1) with
// 3 <= i <= nx-1 and 3 <= j <= ny-1 in strips (every flux and Γ value they
// read is interior), nl_frame_block the ring around them: the rows j = 2
// and ny (i = 2..nx), then the columns i = 2 and nx (j = 3..ny-1).
__host__ __device__ inline int nl_frame_points(int nx, int ny) { return 2 * (nx - 1) + 2 * (ny - 3); }

// The values one output point (i-1, j-1) of nonlinear! reads, as a strip of
// one row (nl_strip with j0 = j1 = j): x-fluxes (i-1+d, j-1+e) at d*2+e,
// y-fluxes (i-1+d, j-1+e) at 6+d*3+e, Γ with its ring (i-2+d, j-2+e) at 12+d*3+e.
constexpr int NLF_V = 21;
struct NLFromLds {
    const GridDev* g;
    const double* v;  // the point's NLF_V values
    int i, j;
    __device__ __forceinline__ double qx(int ii, int jj) const { return v[(ii - i + 1) * 2 + (jj - j + 1)]; }
    __device__ __forceinline__ double qy(int ii, int jj) const { return v[6 + (ii - i + 1) * 3 + (jj - j + 1)]; }
    __device__ __forceinline__ double gam(int ii, int jj) const { return v[12 + (ii - i + 2) * 3 + (jj - j + 2)]; }
    __device__ __forceinline__ double gb(int, int) const { return 0.0; }
};

// One work-group per 64 consecutive ring points of one entry (blockIdx.y).
// Phase 1: wavefront w forms value kinds w, w+4, .. of all 64 points (each
// kind the same curl_qx / curl_qy / ring expression across the wavefront)
// into LDS; phase 2: thread p forms point p from them with nl_strip's
// arithmetic.  Every value is the one the strip kernels form, so the
// result is theirs bit for bit.
// frame block fb of entry b, 256 threads (tid = lane + 64 wave)
__device__ __forceinline__ void nl_frame_block(const NLArgs& a, int fb, int b, int tid)
{
    __shared__ double sv[64][NLF_V + 1];
    const GridDev& g = a.g;
    const int nx = g.nx, ny = g.ny;
    const int e = a.lev0 + b;
    const Entry en = batch_entry(b, a.lev0, a.levmod, 0, a.mg, a.ipar);
    if (!en.active) return;
    const int nf = nl_frame_points(nx, ny);
    auto point = [&](int k, int& i, int& j) -> bool {
        if (k >= nf) return false;
        if (k < 2 * (nx - 1)) {
            i = 2 + k % (nx - 1);
            j = k < nx - 1 ? 2 : ny;
        } else {
            const int c = k - 2 * (nx - 1);
            i = (c & 1) ? nx : 2;
            j = 3 + (c >> 1);
        }
        return true;
    };
    NLLevel lv;
    lv.g = &g;
    lv.G = a.G + (long)e * g.nG;
    lv.Gc = a.G + (long)(e + 1) * g.nG;
    lv.GBl = nullptr;
    lv.hasbc = en.lev < en.L - 1;
    const double* __restrict__ ps = a.psi + (long)e * g.nG;
    const double* __restrict__ psc = a.psi + (long)(e + 1) * g.nG;
    const int lane = tid & (WAVE - 1), w = tid / WAVE;
    int i = 2, j = 2;
    const bool ok = point(fb * WAVE + lane, i, j);
    // wave 0: the six x-fluxes, wave 1: the six y-fluxes, waves 2 and 3: the
    // nine Γ values; each wave's values without branches, so that their loads
    // are in flight together (points past the ring evaluate the point (2, 2))
    if (w == 0) {
#pragma unroll
        for (int kind = 0; kind < 6; kind++)
            sv[lane][kind] = curl_qx_sel(ps, psc, lv.hasbc, g, i - 1 + kind / 2, j - 1 + kind % 2);
    } else if (w == 1) {
#pragma unroll
        for (int kind = 6; kind < 12; kind++)
            sv[lane][kind] = curl_qy_sel(ps, psc, lv.hasbc, g, i - 1 + (kind - 6) / 3, j - 1 + (kind - 6) % 3);
    } else {
        NLFromPsiSel S;
        static_cast<NLLevel&>(S) = lv;
#pragma unroll
        for (int t = 0; t < 5; t++) {
            const int kind = 12 + (w - 2) + 2 * t;  // wave 2: 12, 14, .., 20; wave 3: 13, .., 19
            if (kind < NLF_V) sv[lane][kind] = S.gam(i - 2 + (kind - 12) / 3, j - 2 + (kind - 12) % 3);
        }
    }
    __syncthreads();
    if (w != 0 || !ok) return;
    const double hcl = level_hc(g.h, en.lev);
    double q0vx = a.q0vx, q0vy = a.q0vy;
    if (a.ipar) {
        q0vx = a.ipar[en.inst].q0vx;
        q0vy = a.ipar[en.inst].q0vy;
    }
    NLFromLds S{&g, sv[lane], i, j};
    nl_strip<false, 1>(a, S, a.q0 + (long)e * g.nq, a.q0_uniform != 0, q0vx * hcl, q0vy * hcl, hcl, i, j, j, nullptr,
                       a.out + (long)b * g.nG);
}

template <bool LIN>
__global__ void __launch_bounds__(256) k_nonlinear(NLArgs a)
{
    const GridDev& g = a.g;
    const int nx = g.nx, ny = g.ny;
    // logical block (64 i, 4 strips of NL_JS rows, level), XCD-contiguous
    const int nbx = (nx - 1 + 63) / 64;
    const int nbs = ((ny - 1 + NL_JS - 1) / NL_JS + 3) / 4;
    const int lin = xcd_block(blockIdx.x, gridDim.x / NXCD);
    const int b = lin / (nbx * nbs);
    if (b >= a.nlev) return;
    const int rem = lin - b * (nbx * nbs);
    const int i = (rem % nbx) * 64 + threadIdx.x + 2;                    // 2..nx
    const int j0 = ((rem / nbx) * 4 + threadIdx.y) * NL_JS + 2;          // first output row, 2..ny
    if (i > nx || j0 > ny) return;
    const int j1 = min(j0 + NL_JS - 1, ny);
    const int e = a.lev0 + b;  // batch entry: its fields at e * nq, e * nG
    const Entry en = batch_entry(b, a.lev0, a.levmod, 0, a.mg, a.ipar);
    if (!en.active) return;
    const int lev = en.lev;    // its grid level
    const double* __restrict__ q0 = a.q0 + (long)e * g.nq;
    const bool u0 = a.q0_uniform != 0;
    const double hcl = level_hc(g.h, lev);
    double q0vx = a.q0vx, q0vy = a.q0vy;
    if (a.ipar) {
        q0vx = a.ipar[en.inst].q0vx;
        q0vy = a.ipar[en.inst].q0vy;
    }
    const double cx0 = q0vx * hcl, cy0 = q0vy * hcl;  // as k_base_flux_inertial writes them
    NLLevel lv;
    lv.g = &g;
    lv.G = a.G + (long)e * g.nG;
    lv.Gc = a.G + (long)(e + 1) * g.nG;
    lv.GBl = LIN ? a.GB + (long)lev * g.nG : nullptr;
    lv.hasbc = lev < en.L - 1;
    const double* __restrict__ QBl = LIN ? a.QB + (long)lev * g.nq : nullptr;
    double* __restrict__ out = a.out + (long)b * g.nG;
    NLFromQ S;
    static_cast<NLLevel&>(S) = lv;
    S.q = a.q + (long)e * g.nq;
    nl_strip<LIN, NL_JS>(a, S, q0, u0, cx0, cy0, hcl, i, j0, j1, QBl, out);
}

// nonlinear! with the fluxes from ψ, one launch: the first nint work-groups
// take the interior points in strips of NL_JS rows from j = 3, i = 3..nx-1,
// the rest the ring (nl_frame_block, 64 points each, nfb per entry).
__global__ void __launch_bounds__(256) k_nonlinear_psi(NLArgs a, int nint, int nfb)
{
    const GridDev& g = a.g;
    const int nx = g.nx, ny = g.ny;
    if ((int)blockIdx.x >= nint) {
        const int k = blockIdx.x - nint;
        nl_frame_block(a, k % nfb, k / nfb, threadIdx.x + WAVE * threadIdx.y);
        return;
    }
    const int nbx = (nx - 3 + 63) / 64;
    const int nbs = ((ny - 3 + NL_JS - 1) / NL_JS + 3) / 4;
    const int lin = xcd_block(blockIdx.x, nint / NXCD);
    const int b = lin / (nbx * nbs);
    if (b >= a.nlev) return;
    const int rem = lin - b * (nbx * nbs);
    const int i = (rem % nbx) * 64 + threadIdx.x + 3;                    // 3..nx-1
    const int j0 = ((rem / nbx) * 4 + threadIdx.y) * NL_JS + 3;          // first output row, 3..ny-1
    if (i > nx - 1 || j0 > ny - 1) return;
    const int j1 = min(j0 + NL_JS - 1, ny - 1);
    const int e = a.lev0 + b;
    const Entry en = batch_entry(b, a.lev0, a.levmod, 0, a.mg, a.ipar);
    if (!en.active) return;
    const double hcl = level_hc(g.h, en.lev);
    double q0vx = a.q0vx, q0vy = a.q0vy;
    if (a.ipar) {
        q0vx = a.ipar[en.inst].q0vx;
        q0vy = a.ipar[en.inst].q0vy;
    }
    nl_strip_psi_inner(a.psi + (long)e * g.nG, a.G + (long)e * g.nG, a.q0 + (long)e * g.nq, a.q0_uniform != 0,
                       q0vx * hcl, q0vy * hcl, hcl, i, j0, j1, g, a.out + (long)b * g.nG);
}

void launch_nonlinear(const NLArgs& a, hipStream_t s)
{
    dim3 blk(64, 4);  // k_nonlinear derives its tile from this shape
    const int nblocks = ((a.g.nx - 1 + 63) / 64) * (((a.g.ny - 1 + NL_JS - 1) / NL_JS + 3) / 4) * a.nlev;
    if (a.psi && !a.QB) {
        const int nb = ((a.g.nx - 3 + 63) / 64) * (((a.g.ny - 3 + NL_JS - 1) / NL_JS + 3) / 4) * a.nlev;
        const int nint = NXCD * xcd_nper(nb);
        const int nfb = (nl_frame_points(a.g.nx, a.g.ny) + WAVE - 1) / WAVE;
        hipLaunchKernelGGL(k_nonlinear_psi, dim3(nint + nfb * a.nlev), blk, 0, s, a, nint, nfb);
    } else {
        const dim3 grid(NXCD * xcd_nper(nblocks));
        if (a.QB) hipLaunchKernelGGL(k_nonlinear<true>, grid, blk, 0, s, a);
        else hipLaunchKernelGGL(k_nonlinear<false>, grid, blk, 0, s, a);
    }
}

// ===========================================================================
// curl! on the multigrid (curl.jl:81-116): q = C ψ with ψbc = get_bc!(ψ[:, lev+1])
// (zero on the coarsest level).  Thread (i, j), i = 1..nx+1, j = 1..ny+1,
// writes the x-flux (i, j) (j <= ny) and the y-flux (i, j) (i <= nx): both
// read ψ(i-1, j-1), so each ψ row is fetched by one range of blocks only.
// ===========================================================================
__global__ void __launch_bounds__(256) k_curl(CurlArgs a)
{
    const GridDev& g = a.g;
    const int nx = g.nx, ny = g.ny, n1 = g.n1;
    const int nbx = (nx + 1 + 63) / 64, nby = (ny + 1 + 3) / 4;
    const int lin = xcd_block(blockIdx.x, gridDim.x / NXCD);
    const int b = lin / (nbx * nby);
    if (b >= a.nlev) return;
    const int rem = lin - b * (nbx * nby);
    const int i = (rem % nbx) * 64 + threadIdx.x + 1;  // 1..nx+1
    const int j = (rem / nbx) * 4 + threadIdx.y + 1;   // 1..ny+1
    if (i > nx + 1 || j > ny + 1) return;
    // entry e = lev0 + b: level lev of its field; colbatch: fields at level
    // lev0 with levels psl and fields psb apart
    const int e = a.lev0 + b;
    const Entry en = batch_entry(b, a.lev0, a.levmod, a.colbatch, a.mg, a.ipar);
    if (!en.active) return;
    const int lev = en.lev;
    const long pe = a.colbatch ? (long)lev * a.psl + (long)b * a.psb : (long)e * g.nG;
    const long pl1 = a.colbatch ? a.psl : g.nG;
    const double* __restrict__ psi = a.psi + pe;
    const bool hasbc = lev < en.L - 1;
    const double* __restrict__ pc = a.psi + pe + pl1;
    (void)n1;
    double* __restrict__ q = a.q + (long)b * a.qstride;
    if (j <= ny) q[(i - 1) + (long)(nx + 1) * (j - 1)] = curl_qx(psi, pc, hasbc, g, i, j);
    if (i <= nx) q[g.nu + (i - 1) + (long)nx * (j - 1)] = curl_qy(psi, pc, hasbc, g, i, j);
}

void launch_curl(const CurlArgs& a, hipStream_t s)
{
    const int nblocks = ((a.g.nx + 1 + 63) / 64) * ((a.g.ny + 1 + 3) / 4) * a.nlev;
    hipLaunchKernelGGL(k_curl, dim3(NXCD * xcd_nper(nblocks)), dim3(64, 4), 0, s, a);
}

// ===========================================================================
// coarsify! (mg-utils.jl:71-106): full-weighting restriction into the central
// block of the next coarser level.  One thread per coarse point.
// ===========================================================================
__global__ void __launch_bounds__(256) k_coarsify(CoarsifyArgs a)
{
    const int n1 = a.g.n1;
    // logical block (64 x 4 coarse points), XCD-contiguous (shared fine rows in one L2)
    const int nbx = (a.ci + 63) / 64, nby = (a.cj + 3) / 4;
    const int nbt = a.nb > 1 ? a.nb : 1;
    const int lin0 = xcd_block(blockIdx.x, gridDim.x / NXCD);
    const int b = lin0 / (nbx * nby);
    if (b >= nbt) return;
    if (a.ipar && a.clev > a.ipar[b].L) return;  // instance b has no level clev
    const int lin = lin0 - b * (nbx * nby);
    const int ai = (lin % nbx) * 64 + threadIdx.x;
    const int bj = (lin / nbx) * 4 + threadIdx.y;
    if (ai >= a.ci || bj >= a.cj) return;
    const int i = a.i0 + 2 * ai, ic = a.ic0 + ai;
    const int j = a.j0 + 2 * bj, jc = a.jc0 + bj;
    const double* __restrict__ G = a.G + (long)b * a.bstride;
    auto F = [&](int ii, int jj) -> double { return G[(ii - 1) + (long)n1 * (jj - 1)]; };
    double s1 = ((F(i + 1, j) + F(i, j + 1)) + F(i - 1, j)) + F(i, j - 1);
    double s2 = ((F(i + 1, j + 1) + F(i + 1, j - 1)) + F(i - 1, j - 1)) + F(i - 1, j + 1);
    a.Gc[(long)b * a.bstride + (ic - 1) + (long)n1 * (jc - 1)] = (F(i, j) + 0.5 * s1) + 0.25 * s2;
}

void launch_coarsify(const CoarsifyArgs& a, hipStream_t s)
{
    if (a.ci <= 0 || a.cj <= 0) return;
    const int nblocks = ((a.ci + 63) / 64) * ((a.cj + 3) / 4) * (a.nb > 1 ? a.nb : 1);
    hipLaunchKernelGGL(k_coarsify, dim3(NXCD * xcd_nper(nblocks)), dim3(64, 4), 0, s, a);
}

// ===========================================================================
// rot! (curl.jl:22-34) over a rectangle of output points
// ===========================================================================
__global__ void __launch_bounds__(256) k_rot(const double* __restrict__ q, double* __restrict__ G, int ib0, int ib1,
                                             int jb0, int jb1, GridDev g, long qstride, long gstride)
{
    const int i = ib0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = jb0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > ib1 || j > jb1) return;
    q += (long)blockIdx.z * qstride;
    G += (long)blockIdx.z * gstride;
    const int nx = g.nx;
    auto QXv = [&](int ii, int jj) -> double { return q[(ii - 1) + (long)(nx + 1) * (jj - 1)]; };
    auto QYv = [&](int ii, int jj) -> double { return q[g.nu + (ii - 1) + (long)nx * (jj - 1)]; };
    double a = QXv(i, j - 1) - QXv(i, j);
    double c = QYv(i, j) - QYv(i - 1, j);
    G[(i - 2) + (long)g.n1 * (j - 2)] = a + c;
}

void launch_rot(const double* q, double* G, int ib0, int ib1, int jb0, int jb1, const GridDev& g, hipStream_t s)
{
    if (ib1 < ib0 || jb1 < jb0) return;
    dim3 blk(64, 4);
    dim3 grid((ib1 - ib0 + 1 + 63) / 64, (jb1 - jb0 + 1 + 3) / 4);
    hipLaunchKernelGGL(k_rot, grid, blk, 0, s, q, G, ib0, ib1, jb0, jb1, g, 0L, 0L);
}

void launch_rot_batch(const double* q, long qstride, double* G, long gstride, int nb, int ib0, int ib1, int jb0,
                      int jb1, const GridDev& g, hipStream_t s)
{
    if (ib1 < ib0 || jb1 < jb0 || nb <= 0) return;
    dim3 blk(64, 4);
    dim3 grid((ib1 - ib0 + 1 + 63) / 64, (jb1 - jb0 + 1 + 3) / 4, nb);
    hipLaunchKernelGGL(k_rot, grid, blk, 0, s, q, G, ib0, ib1, jb0, jb1, g, qstride, gstride);
}

// ===========================================================================
// E = regT! (interface-coupling.jl:122-133): one wavefront per (point, component)
// over its 13 x 13 window; fb = sum(weight .* (q + q0)) (boundary_forces!
// forms q + q0 first, timestepping.jl:158-159).
// ===========================================================================
__global__ void __launch_bounds__(256) k_gather(GatherArgs a)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid0 = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int bq = wid0 / (2 * a.nbt);
    if (bq >= (a.nb > 1 ? a.nb : 1)) return;
    const int wid = wid0 - bq * (2 * a.nbt);
    const int c = wid / a.nbt, k = wid % a.nbt;
    const double* __restrict__ qf = a.q + (long)bq * a.qstride;
    const double* __restrict__ q0f = a.q0;
    const double* __restrict__ ps = a.psi;
    const double* __restrict__ pc = a.psic;
    bool hasbc = a.psic != nullptr;
    long o = (long)bq * a.ostride + c * a.nbt + k;
    if (a.pinst) {  // a group: point k is point plocal[k] of instance pinst[k]
        const int inst = a.pinst[k];
        qf = a.q + (long)inst * a.qstride;
        if (q0f) q0f += (long)inst * a.q0stride;
        o = a.ifoff[inst] + (long)c * a.inbt[inst] + a.plocal[k];
        if (ps) {
            ps = a.psi + (long)inst * a.psistride;
            pc = ps + a.g.nG;
            hasbc = hasbc && (!a.ipar || a.ipar[inst].L > 1);
        }
    }
    const int bx = a.bidx[k], by = a.bidx[a.nbt + k];
    const double* w = a.w + (long)(2 * k + c) * 169;
    const int nx = a.g.nx;
    double acc = 0.0;
    for (int e = lane; e < 169; e += WAVE) {
        const int l = e % 13, m = e / 13;
        const int i = bx + l - 6, j = by + m - 6;
        long p = (c == 0) ? (i - 1) + (long)(nx + 1) * (j - 1) : a.g.nu + (i - 1) + (long)nx * (j - 1);
        double f;
        if (ps) f = (c == 0) ? curl_qx(ps, pc, hasbc, a.g, i, j) : curl_qy(ps, pc, hasbc, a.g, i, j);
        else f = qf[p];
        double v = q0f ? (f + q0f[p]) : f;
        acc += w[e] * v;
    }
    acc = wave_sum(acc);
    if (lane == 0) a.out[o] = acc;
}

void launch_gather(const GatherArgs& a, hipStream_t s)
{
    long waves = 2L * a.nbt * (a.nb > 1 ? a.nb : 1);
    if (a.nbt == 0) return;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, a);
}

// E' e_f: field b gets the weights of F̃b entry f = f0 + b at its window cells
// (the only non-zero terms of reg! for a unit force, interface-coupling.jl:99-119)
__global__ void __launch_bounds__(256) k_unit_reg(double* q, long qstride, int nb, int f0, const int* bidx,
                                                  const double* w, int nbt, GridDev g)
{
    const int b = blockIdx.x;
    if (b >= nb) return;
    const int f = f0 + b;
    const int c = f / nbt, k = f - c * nbt;
    const int bx = bidx[k], by = bidx[nbt + k];
    const double* wk = w + (long)(2 * k + c) * 169;
    double* qb = q + (long)b * qstride;
    for (int e = threadIdx.x; e < 169; e += blockDim.x) {
        const int l = e % 13, m = e / 13;
        const int i = bx + l - 6, j = by + m - 6;
        long p = (c == 0) ? (i - 1) + (long)(g.nx + 1) * (j - 1) : g.nu + (i - 1) + (long)g.nx * (j - 1);
        qb[p] = wk[e];
    }
}

void launch_unit_reg(double* q, long qstride, int nb, int f0, const int* bidx, const double* w, int nbt,
                     const GridDev& g, hipStream_t s)
{
    if (nb <= 0 || nbt == 0) return;
    hipLaunchKernelGGL(k_unit_reg, dim3(nb), dim3(256), 0, s, q, qstride, nb, f0, bidx, w, nbt, g);
}

// ===========================================================================
// E' = reg! (interface-coupling.jl:99-119) as an ordered segmented gather:
// each touched flux cell sums its contributions in increasing body-point
// order, which is the order of the reference's serial accumulation.
// ===========================================================================
__global__ void __launch_bounds__(256) k_scatter(ScatterArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncell) return;
    double acc = 0.0;
    for (int e = a.off[t]; e < a.off[t + 1]; e++) acc = acc + a.ent_w[e] * a.fb[a.ent_f[e]];
    a.q[a.cell[t]] = acc;
}

void launch_scatter(const ScatterArgs& a, hipStream_t s)
{
    if (a.ncell == 0) return;
    hipLaunchKernelGGL(k_scatter, dim3((a.ncell + 255) / 256), dim3(256), 0, s, a);
}

// ===========================================================================
// F̃b = Binv (E (q + q0)) (boundary_forces!, timestepping.jl:160), then
// update_stress! (timestepping.jl:330-354) in a second, one-block launch.
// (A single launch whose last work-group forms the forces needs an
// agent-scope release per work-group -- an L2 writeback -- and measured
// slower than the extra launch.)
// ===========================================================================
__device__ __forceinline__ void forces_block(const ForceArgs& a)
{
    __shared__ double red[2][4];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
    const long ho = a.slot ? (long)(*a.slot) * a.nbh : 0;  // history entry of this step
    for (int bd = 0; bd < a.nbodies; bd++) {
        const int o = a.boff[bd], n = a.boff[bd + 1] - o;
        double sx = 0.0, sy = 0.0;
        for (int k = tid; k < n; k += blockDim.x) {
            double ds = a.ds[o + k];
            double fx = (a.Fb[o + k] * a.hdt) / ds;
            double fy = (a.Fb[a.nbt + o + k] * a.hdt) / ds;
            a.fb[2L * o + k] = fx;
            a.fb[2L * o + n + k] = fy;
            sx += ds * fx;
            sy += ds * fy;
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        if (lane == 0) { red[0][w] = sx; red[1][w] = sy; }
        __syncthreads();
        if (tid == 0) {
            double X = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
            double Y = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
            a.CD[bd] = 2.0 * X;
            a.CL[bd] = 2.0 * Y;
            if (a.cd_hist) { a.cd_hist[ho + bd] = 2.0 * X; a.cl_hist[ho + bd] = 2.0 * Y; }
        }
        __syncthreads();
    }
    if (a.slot && tid == 0) *a.slot = (int)(ho / a.nbh) + 1;
}

// GEMV: one wavefront per row of Binv (given transposed: AT row r = row r of
// Binv), each lane summing its columns in increasing order
__global__ void __launch_bounds__(256) k_gemv_rows(const double* __restrict__ AT, const double* __restrict__ x,
                                                   double* __restrict__ y, int n)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (r >= n) return;
    const double* row = AT + (long)r * n;
    double acc = 0.0;
    // loads of a chunk in flight before its (in-order) fma chain
    constexpr int U = 8;
    for (int c0 = 0; c0 < n; c0 += U * WAVE) {
        double av[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + lane + u * WAVE;
            av[u] = c < n ? row[c] : 0.0;
            xv[u] = c < n ? x[c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (c0 + lane + u * WAVE < n) acc = fma(av[u], xv[u], acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) y[r] = acc;
}

// update_stress!: fb = F̃b (h/dt) ./ ds, CD = 2 sum(ds .* fb[:,1]),
// CL = 2 sum(ds .* fb[:,2]) per body; one block
__global__ void __launch_bounds__(256) k_forces(ForceArgs fa) { forces_block(fa); }

// E' F̃b (k_scatter) and update_stress! (forces_block) in one launch: both only
// read F̃b; the last work-group forms the forces
__global__ void __launch_bounds__(256) k_scatter_forces(ScatterArgs a, ForceArgs fa, int nsb)
{
    if ((int)blockIdx.x == nsb) {
        forces_block(fa);
        return;
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncell) return;
    double acc = 0.0;
    for (int e = a.off[t]; e < a.off[t + 1]; e++) acc = acc + a.ent_w[e] * a.fb[a.ent_f[e]];
    a.q[a.cell[t]] = acc;
}

void launch_scatter_forces(const ScatterArgs& a, const ForceArgs& fa, hipStream_t s)
{
    const int nsb = (a.ncell + 255) / 256;
    if (fa.nbodies == 0) {
        launch_scatter(a, s);
        return;
    }
    hipLaunchKernelGGL(k_scatter_forces, dim3(nsb + 1), dim3(256), 0, s, a, fa, nsb);
}

void launch_gemv(const double* AT, const double* x, double* y, int n, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_gemv_rows, dim3((n + 3) / 4), dim3(256), 0, s, AT, x, y, n);
}

void launch_gemv_forces(const double* AT, const double* x, double* y, int n, const ForceArgs& fa, hipStream_t s)
{
    if (n == 0 || fa.nbodies == 0) return;
    hipLaunchKernelGGL(k_gemv_rows, dim3((n + 3) / 4), dim3(256), 0, s, AT, x, y, n);
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(256), 0, s, fa);
}

void launch_forces(const ForceArgs& fa, hipStream_t s)
{
    if (fa.nbodies == 0) return;
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(256), 0, s, fa);
}

// a group's GEMVs: row r of all instances' rows is row rloc[r] of instance
// rinst[r]; the same wavefront-per-row sum as k_gemv_rows
__global__ void __launch_bounds__(256) k_gemv_rows_group(const double* __restrict__ AT, const long* __restrict__ aoff,
                                                         const double* __restrict__ x, double* __restrict__ y,
                                                         const int* __restrict__ rinst, const int* __restrict__ rloc,
                                                         const int* __restrict__ ifoff, const int* __restrict__ inf,
                                                         int nrows)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int rr = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    if (rr >= nrows) return;
    const int inst = rinst[rr], r = rloc[rr], n = inf[inst];
    const double* row = AT + aoff[inst] + (long)r * n;
    const double* xi = x + ifoff[inst];
    double acc = 0.0;
    constexpr int U = 8;
    for (int c0 = 0; c0 < n; c0 += U * WAVE) {
        double av[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + lane + u * WAVE;
            av[u] = c < n ? row[c] : 0.0;
            xv[u] = c < n ? xi[c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (c0 + lane + u * WAVE < n) acc = fma(av[u], xv[u], acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) y[ifoff[inst] + r] = acc;
}

// update_stress! of every instance of a group: block i takes instance i
__global__ void __launch_bounds__(256) k_forces_group(const ForceArgs* __restrict__ fa) { forces_block(fa[blockIdx.x]); }

void launch_gemv_forces_group(const double* AT, const long* aoff, const double* x, double* y, const int* rinst,
                              const int* rloc, const int* ifoff, const int* inf, int nrows, const ForceArgs* fa,
                              int n, hipStream_t s)
{
    if (nrows == 0 || n == 0) return;
    hipLaunchKernelGGL(k_gemv_rows_group, dim3((nrows + 3) / 4), dim3(256), 0, s, AT, aoff, x, y, rinst, rloc, ifoff,
                       inf, nrows);
    hipLaunchKernelGGL(k_forces_group, dim3(n), dim3(256), 0, s, fa);
}

// ===========================================================================
// base_flux! for MovingGrid (timestepping.jl:273-318) with XX, YY of
// models.jl:140-153 (xc = yc = 0) generated from indices.
// ===========================================================================
__global__ void __launch_bounds__(256) k_base_flux_moving(BaseFluxArgs a)
{
    const GridDev& g = a.g;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lev = blockIdx.y;
    if (p >= g.nq) return;
    const double hc = g.h * (double)(1L << lev);
    double v;
    if (p < g.nu) {
        const int j = (int)(p / (g.nx + 1)) + 1;
        double y = ((((double)j - 0.5) - (double)g.ny / 2.0) * hc + ((double)g.ny / 2.0) * g.h) - a.offy;
        double yy = y - 0.0;
        v = yy * (-hc * a.Om);
        v += hc * a.Ux0;
    } else {
        const long pp = p - g.nu;
        const int i = (int)(pp % g.nx) + 1;
        double x = ((((double)i - 0.5) - (double)g.nx / 2.0) * hc + ((double)g.nx / 2.0) * g.h) - a.offx;
        double xx = x - 0.0;
        v = xx * (hc * a.Om);
        v += hc * a.Uy0;
    }
    a.q0[(long)lev * g.nq + p] = v;
}

void launch_base_flux_moving(const BaseFluxArgs& a, hipStream_t s)
{
    dim3 grid((unsigned)((a.g.nq + 255) / 256), a.mg);
    hipLaunchKernelGGL(k_base_flux_moving, grid, dim3(256), 0, s, a);
}

// base_flux! for InertialMotion (timestepping.jl:252-270): uniform per level.
__global__ void __launch_bounds__(256) k_base_flux_inertial(double* q0, double vx, double vy, GridDev g)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lev = blockIdx.y;
    if (p >= g.nq) return;
    const double hc = g.h * (double)(1L << lev);
    q0[(long)lev * g.nq + p] = (p < g.nu) ? vx * hc : vy * hc;
}

void launch_base_flux_inertial(double* q0, int mg, double vx, double vy, const GridDev& g, hipStream_t s)
{
    dim3 grid((unsigned)((g.nq + 255) / 256), mg);
    hipLaunchKernelGGL(k_base_flux_inertial, grid, dim3(256), 0, s, q0, vx, vy, g);
}

// ===========================================================================
// small utilities
// ===========================================================================
__global__ void k_fill(double* p, long n, double v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_set_int(int* p, int v) { *p = v; }
void launch_set_int(int* p, int v, hipStream_t s) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, v); }

void launch_fill(double* p, long n, double v, hipStream_t s)
{
    if (n <= 0) return;
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

__global__ void k_copy(double* d, const double* src, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) d[i] = src[i];
}
void launch_copy(double* dst, const double* src, long n, hipStream_t s)
{
    if (n <= 0) return;
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
}

__global__ void k_set_unit(double* p, int n, int j)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = (i == j) ? 1.0 : 0.0;
}
void launch_set_unit(double* p, int n, int j, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(256), 0, s, p, n, j);
}

// max that keeps a NaN (fmax drops it): maximum(abs.(q)) of a state holding a
// NaN is NaN in the reference (ibpm.jl:83-88)
__device__ inline double nanmax(double m, double v)
{
    if (m != m) return m;
    if (v != v) return v;
    return v > m ? v : m;
}

__global__ void __launch_bounds__(256) k_absmax(const double* p, long n, double* out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) m = nanmax(m, fabs(p[i]));
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off, WAVE));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = nanmax(nanmax(red[0], red[1]), nanmax(red[2], red[3]));
}
void launch_absmax(const double* p, long n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_absmax, dim3(1), dim3(256), 0, s, p, n, out);
}

void kernels_init() {}

}  // namespace ibpm
