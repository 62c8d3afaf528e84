// ibpm_dev.h -- device-side building blocks shared by the HIP kernels.
//
// Layouts follow the reference exactly (Julia column-major, 1-based in the
// comments): a circulation-like field of one grid level is an (nx-1) x (ny-1)
// array; a flux field of one level is the x-fluxes as an (nx+1) x ny array
// followed by the y-fluxes as an nx x (ny+1) array
// (src/fluid-domain/fluid-domain.jl:55-65).  Levels are consecutive columns
// of an (n x mg) matrix.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ibpm {

constexpr int WAVE = 64;
// Largest DST line the wave-level transform handles: M = n + 1 <= MAXM
// (half-length FFT of P = M/2 <= MAXM/2 points).  Register staging arrays are
// sized from it.
constexpr int MAXM = 1024;
constexpr int MAXP = MAXM / 2;

// Grid geometry passed by value to every kernel (MultiGrid,
// src/fluid-domain/fluid-domain.jl:10-25).
struct GridDev {
    int nx, ny;        // cells
    int n1, n2;        // nx-1, ny-1
    long nG;           // (nx-1)(ny-1)
    long nu, nv, nq;   // ny(nx+1), nx(ny+1), nu+nv
    int LEFT, RIGHT, BOT, TOP, nbc;
    double h;
};

// DST-I plan for lines of length n = M-1 (FFTW RODFT00, unnormalised).
// kind 0: half-length complex FFT (P = M/2 points, Stockham, radices in
// `radix`); kind 1: direct O(M^2) sum (odd M or unsupported factors).
struct DstPlan {
    int M;
    int P;
    int kind;
    int nst;
    unsigned rad;       // radix of stage k in bits 4k..4k+3 (at most 8 stages)
    const double2* tw;  // exp(-2 pi i e / P), e < P
    const double* sn;   // kind 0: sin(pi j / M), j <= M/2;  kind 1: sin(pi t / M), t < 2M
    const double2* wm;  // exp(-2 pi i k / M), k < P
};

__host__ __device__ inline int plan_radix(const DstPlan& pl, int k) { return (int)((pl.rad >> (4 * k)) & 15u); }

// grid spacing of level lev (0-based): h 2^lev, exact
__host__ __device__ inline double level_hc(double h, int lev) { return h * (double)(1L << lev); }

// gfx950 deals the work-groups of a launch round-robin over its 8 XCDs, each
// with a private L2.  A 1-D launch of NXCD * nper work-groups that reads its
// work-group index through xcd_block() hands each XCD a contiguous range of
// nper logical blocks, so neighbouring tiles that share cache lines are
// fetched into one L2 instead of up to eight.
constexpr int NXCD = 8;
__host__ __device__ inline int xcd_nper(int nblocks) { return (nblocks + NXCD - 1) / NXCD; }
__device__ __forceinline__ int xcd_block(int bid, int nper) { return (bid % NXCD) * nper + bid / NXCD; }

// ---------------------------------------------------------------------------
// wave-level synchronisation of LDS traffic (all lanes of one wavefront)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ __forceinline__ double2 cmi(double2 a) { return make_double2(a.y, -a.x); }

// ---------------------------------------------------------------------------
// small forward DFTs (W_R = exp(-2 pi i / R)) on registers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dft2(double2* v)
{
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

__device__ __forceinline__ void dft4(double2* v)
{
    double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    double2 s13 = cadd(v[1], v[3]), d13 = cmi(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

__device__ __forceinline__ void dft3(double2* v)
{
    const double c = -0.5, s = 0.86602540378443864676;  // cos(2pi/3), sin(2pi/3)
    double2 t = cadd(v[1], v[2]);
    double2 d = csub(v[1], v[2]);
    double2 m = make_double2(fma(c, t.x, v[0].x), fma(c, t.y, v[0].y));
    // -i * s * d
    double2 r = make_double2(s * d.y, -s * d.x);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

__device__ __forceinline__ void dft5(double2* v)
{
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2pi/5), cos(4pi/5)
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2pi/5), sin(4pi/5)
    double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    double2 x0 = v[0];
    double2 m1 = make_double2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    double2 m2 = make_double2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    // -i * (s1 b1 + s2 b2), -i * (s2 b1 - s1 b2)
    double2 n1 = make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    double2 n2 = make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    n1 = cmi(n1);
    n2 = cmi(n2);
    v[0] = make_double2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

__device__ __forceinline__ void dft8(double2* v)
{
    const double r = 0.70710678118654752440;
    // radix-2 over stride 4
    double2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
    double2 a1 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
    double2 a2 = cadd(v[2], v[6]), a6 = csub(v[2], v[6]);
    double2 a3 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
    // twiddles W8^1, W8^2, W8^3 on the odd half
    a5 = make_double2(r * (a5.x + a5.y), r * (a5.y - a5.x));     // * (1 - i)/sqrt2
    a6 = cmi(a6);                                                // * -i
    a7 = make_double2(r * (a7.y - a7.x), -r * (a7.x + a7.y));    // * (-1 - i)/sqrt2
    double2 e[4] = {a0, a1, a2, a3};
    double2 o[4] = {a4, a5, a6, a7};
    dft4(e);
    dft4(o);
    v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
    v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
}

// The complex work array of the FFT is padded by one element per 8: element e
// lives at pz(e).  Stockham writes of the first stages (stride R or R*Ns
// complex) then fall in distinct LDS banks instead of one.
__device__ __forceinline__ int pz(int e) { return e + (e >> 3); }

// 10-point DFT as 2 x 5 by the prime-factor (Good-Thomas) index map, so
// that no twiddles enter between the two sets of small DFTs: input
// n = (5 n1 + 2 n2) mod 10, output k = (5 k1 + 6 k2) mod 10, and
// W10^{nk} = W2^{n1 k1} W5^{n2 k2}.
__device__ __forceinline__ void dft10(double2* v)
{
    double2 a[2][5];
#pragma unroll
    for (int n2 = 0; n2 < 5; n2++) {
        const double2 x0 = v[(2 * n2) % 10], x1 = v[(5 + 2 * n2) % 10];
        a[0][n2] = cadd(x0, x1);
        a[1][n2] = csub(x0, x1);
    }
    dft5(a[0]);
    dft5(a[1]);
#pragma unroll
    for (int k1 = 0; k1 < 2; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 5; k2++) v[(5 * k1 + 6 * k2) % 10] = a[k1][k2];
}

template <int R>
__device__ __forceinline__ void dft_r(double2* v)
{
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 5) dft5(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 10) dft10(v);
}

// One in-place Stockham stage of radix R over P complex points in LDS `z`
// (padded), executed by one wavefront.  Ns = product of the earlier radices.
template <int R>
__device__ __forceinline__ void fft_stage(double2* z, int P, int Ns, const double2* __restrict__ tw, int lane)
{
    constexpr int TMAX = (MAXP / R + WAVE - 1) / WAVE;
    const int nb = P / R;
    const int tws = P / (Ns * R);
    double2 v[TMAX][R];
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        int j = lane + WAVE * t;
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = z[pz(j + r * nb)];
        }
    }
    wave_sync();
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        int j = lane + WAVE * t;
        if (j < nb) {
            int k = j % Ns;
#pragma unroll
            for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], tw[k * r * tws]);
            dft_r<R>(v[t]);
            int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) z[pz(base + r * Ns)] = v[t][r];
        }
    }
    wave_sync();
}

// y_j = sin(pi j/M)(x_j + x_{M-j}) + (x_j - x_{M-j})/2 with x_j = s[j-1],
// x_0 = x_M = 0 and sin(pi j/M) = sn[min(j, M-j)]
__device__ __forceinline__ double pre_y(const double* s, int M, const double* sn, int j)
{
    if (j == 0) return 0.0;
    const int jm = M - j;
    const double xa = s[j - 1];
    const double xb = (jm >= 1) ? s[jm - 1] : 0.0;
    const double sj = sn[j < jm ? j : jm];
    return fma(sj, xa + xb, 0.5 * (xa - xb));
}

// First Stockham stage (Ns = 1, no twiddles) fed directly by the
// pre-processing: butterfly j takes z_m = y_{2m} + i y_{2m+1} for
// m = j + r P/R.  All reads of the real input s precede all writes of z.
template <int R>
__device__ __forceinline__ void fft_first_stage(double* s, double2* z, int P, int M, const double* __restrict__ sn,
                                                int lane)
{
    constexpr int RR = R > 1 ? R : 1;
    constexpr int TMAX = (MAXP / RR + WAVE - 1) / WAVE;
    const int nb = P / RR;
    double2 v[TMAX][RR];
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        int j = lane + WAVE * t;
        if (j < nb) {
#pragma unroll
            for (int r = 0; r < RR; r++) {
                int m = j + r * nb;
                v[t][r] = make_double2(pre_y(s, M, sn, 2 * m), pre_y(s, M, sn, 2 * m + 1));
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        int j = lane + WAVE * t;
        if (j < nb) {
            dft_r<RR>(v[t]);
#pragma unroll
            for (int r = 0; r < RR; r++) z[pz(j * RR + r)] = v[t][r];
        }
    }
    wave_sync();
}

// DPP move of a double (two 32-bit halves); lanes the control does not feed
// read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Wave-level inclusive prefix sum of one double per lane with DPP: shifts by
// 1, 2, 4, 8 within rows of 16 lanes, then the row-15 and row-31 broadcasts
// carry the partial sums across rows.  No LDS traffic.
__device__ __forceinline__ double wave_incl_scan(double v, int lane)
{
    (void)lane;
    v += dpp_d<0x111, 0xf>(v);  // row_shr:1
    v += dpp_d<0x112, 0xf>(v);  // row_shr:2
    v += dpp_d<0x114, 0xf>(v);  // row_shr:4
    v += dpp_d<0x118, 0xf>(v);  // row_shr:8
    v += dpp_d<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_d<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// value of lane 63 in every lane
__device__ __forceinline__ double wave_last(double v)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), WAVE - 1);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), WAVE - 1);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// ---------------------------------------------------------------------------
// In-place DST-I of one line held in LDS, executed by one wavefront.
//   on entry: s[n] = X_n (n = 0..M-2), s has room for M doubles (16B aligned)
//   on exit : s[n] = Y_n = 2 sum_j X_j sin(pi (j+1)(n+1)/M)
// Algorithm (kind 0, M even, P = M/2): with x_j = X_{j-1}, x_0 = x_M = 0,
//   y_j = sin(pi j/M)(x_j + x_{M-j}) + (x_j - x_{M-j})/2      (j = 0..M-1)
//   F = real DFT of y via the P-point complex FFT of z_m = y_{2m} + i y_{2m+1};
//   with F_k = R_k - i I_k:  Y(2k) = 2 I_k,  Y(2k+1) = R_0 + 2 sum_{m=1..k} R_m
//   (Y(p) = 2 sum_j x_j sin(pi j p / M)).  Both identities follow from the
//   symmetric/antisymmetric split of y under j -> M-j.
// ---------------------------------------------------------------------------
constexpr int MAX_PAIRS = (MAXP + WAVE - 1) / WAVE;

// The plan's tables staged in LDS: the FFT stages, the pre-processing and the
// unpacking read them at every butterfly, so they must not cost an L2 round
// trip each time.
struct DstTab {
    const double2* tw;
    const double* sn;
    const double2* wm;
};

// doubles taken by the LDS copy of a plan's tables (kept even: 16 B aligned)
__host__ __device__ inline int dst_tab_doubles(const DstPlan& pl)
{
    return pl.kind == 0 ? 4 * pl.P + ((pl.P + 2) & ~1) : 2 * pl.M;
}

// doubles of the LDS slot one line is transformed in: the real line (M-1
// values), the padded complex work array (pz(P-1) + 1 elements) and the
// linear unpacking arrays (2P = M values) all live in it; +2 staggers the
// banks of neighbouring slots.
__host__ __device__ inline int dst_slot_doubles(const DstPlan& pl)
{
    int n = pl.M;
    if (pl.kind == 0) {
        int c = 2 * (pl.P + (pl.P - 1) / 8);
        if (c > n) n = c;
    }
    return ((n + 1) & ~1) + 2;
}

// Cooperative copy of the tables into `lds` by threads tid = 0..nthreads-1;
// the caller synchronises the workgroup before use.
__device__ inline DstTab load_dst_tab(double* lds, const DstPlan& pl, int tid, int nthreads)
{
    DstTab t;
    if (pl.kind == 0) {
        double2* tw = reinterpret_cast<double2*>(lds);
        double2* wm = tw + pl.P;
        double* sn = reinterpret_cast<double*>(wm + pl.P);
        for (int i = tid; i < pl.P; i += nthreads) {
            tw[i] = pl.tw[i];
            wm[i] = pl.wm[i];
        }
        for (int i = tid; i <= pl.P; i += nthreads) sn[i] = pl.sn[i];
        t.tw = tw;
        t.wm = wm;
        t.sn = sn;
    } else {
        for (int i = tid; i < 2 * pl.M; i += nthreads) lds[i] = pl.sn[i];
        t.tw = nullptr;
        t.wm = nullptr;
        t.sn = lds;
    }
    return t;
}

__device__ __forceinline__ void dst1_line(double* s, const DstPlan& pl, const DstTab& tb, int lane)
{
    // Opaque copies of the lane and sizes: the transform's index arithmetic is
    // loop-invariant where the caller loops over lines, and hoisting all of it
    // out of the loop would keep hundreds of addresses live (register spills).
    asm volatile("" : "+v"(lane));
    int M = pl.M;
    asm volatile("" : "+s"(M));
    if (pl.kind == 1) {
        // direct sum: Y_{k} (k = 1..M-1) = 2 sum_j x_j sin(pi j k / M)
        constexpr int OUTMAX = (MAXM + WAVE - 1) / WAVE;
        double acc[OUTMAX];
#pragma unroll
        for (int t = 0; t < OUTMAX; t++) {
            int k = lane + WAVE * t + 1;
            acc[t] = 0.0;
            if (k <= M - 1) {
                double a = 0.0;
                int e = 0;  // (j*k) mod 2M
                for (int j = 1; j <= M - 1; j++) {
                    e += k;
                    if (e >= 2 * M) e -= 2 * M;
                    a = fma(s[j - 1], tb.sn[e], a);
                }
                acc[t] = 2.0 * a;
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < OUTMAX; t++) {
            int k = lane + WAVE * t + 1;
            if (k <= M - 1) s[k - 1] = acc[t];
        }
        wave_sync();
        return;
    }
    int P = pl.P;
    asm volatile("" : "+s"(P));
    double2* z = reinterpret_cast<double2*>(s);
    // --- pre-processing fused into the first FFT stage (reads X, writes the
    //     stage output into the padded complex array)
    {
        const int R = plan_radix(pl, 0);
        if (R == 8) fft_first_stage<8>(s, z, P, M, tb.sn, lane);
        else if (R == 4) fft_first_stage<4>(s, z, P, M, tb.sn, lane);
        else if (R == 2) fft_first_stage<2>(s, z, P, M, tb.sn, lane);
        else if (R == 5) fft_first_stage<5>(s, z, P, M, tb.sn, lane);
        else if (R == 3) fft_first_stage<3>(s, z, P, M, tb.sn, lane);
        else fft_first_stage<1>(s, z, P, M, tb.sn, lane);  // P = 1: no stage
    }
    // --- remaining Stockham stages (natural order out)
    int Ns = pl.nst > 0 ? plan_radix(pl, 0) : 1;
    for (int st = 1; st < pl.nst; st++) {
        int R = plan_radix(pl, st);
        if (R == 8) fft_stage<8>(z, P, Ns, tb.tw, lane);
        else if (R == 4) fft_stage<4>(z, P, Ns, tb.tw, lane);
        else if (R == 2) fft_stage<2>(z, P, Ns, tb.tw, lane);
        else if (R == 5) fft_stage<5>(z, P, Ns, tb.tw, lane);
        else fft_stage<3>(z, P, Ns, tb.tw, lane);
        Ns *= R;
    }
    // --- real-DFT unpacking for pairs (k, P-k), k = 0..P/2, into linear arrays
    //     R'_k at s[k] (R'_0 = R_0, R'_k = 2 R_k) and Y(2k) = 2 I_k at s[P + k]
    {
        constexpr int KMAX = (MAXP / 2 + 1 + WAVE - 1) / WAVE;
        double2 za[KMAX], zb[KMAX];
        const int nk = P / 2 + 1;
#pragma unroll
        for (int t = 0; t < KMAX; t++) {
            int k = lane + WAVE * t;
            if (k < nk) {
                int k2 = (P - k) % P;
                za[t] = z[pz(k)];
                zb[t] = z[pz(k2)];
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < KMAX; t++) {
            int k = lane + WAVE * t;
            if (k < nk) {
                int k2 = (P - k) % P;
                double2 a = za[t], b = zb[t];
                // F_k = E + W^k O,  E = (a + conj b)/2,  O = -i (a - conj b)/2
                {
                    double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
                    double2 O = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
                    double2 F = cadd(E, cmul(tb.wm[k], O));
                    s[k] = (k == 0) ? F.x : 2.0 * F.x;
                    if (k > 0) s[P + k] = -2.0 * F.y;
                }
                if (k2 != k) {
                    double2 E = make_double2(0.5 * (b.x + a.x), 0.5 * (b.y - a.y));
                    double2 O = make_double2(0.5 * (b.y + a.y), -0.5 * (b.x - a.x));
                    double2 F = cadd(E, cmul(tb.wm[k2], O));
                    s[k2] = 2.0 * F.x;
                    s[P + k2] = -2.0 * F.y;
                }
            }
        }
        wave_sync();
    }
    // --- prefix sum of R' in rounds of 64 consecutive k: Y(2k+1) = sum_{m<=k} R'_m;
    //     then both output halves are written in natural order, s[n] = Y(n+1)
    {
        constexpr int NR = (MAXP + WAVE - 1) / WAVE;
        double sc[NR], ev[NR];
        double carry = 0.0;
#pragma unroll
        for (int t = 0; t < NR; t++) {
            int k = WAVE * t + lane;
            if (WAVE * t < P) {
                double v = (k < P) ? s[k] : 0.0;
                double incl = wave_incl_scan(v, lane);
                sc[t] = carry + incl;
                carry += wave_last(incl);
                ev[t] = (k >= 1 && k < P) ? s[P + k] : 0.0;
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < NR; t++) {
            int k = WAVE * t + lane;
            if (k < P) {
                s[2 * k] = sc[t];
                if (k >= 1) s[2 * k - 1] = ev[t];
            }
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// The same transform with P = M/2 known at compile time: the radix plan, the
// trip counts, the Stockham index arithmetic and the padding all fold into
// constants, which removes most of the per-lane integer work of dst1_line.
// ---------------------------------------------------------------------------
struct RadPlan {
    int n;
    int r[8];
};

// radices of P: 8s first, then the remaining 4 or 2, then 5s and 3s, as the
// host plan, except that a remaining 2 takes a 5 along as one radix-10 stage
// where that stage still has half a wavefront of butterflies (one LDS round
// trip fewer; with fewer lanes busy, as for P = 150, the merged stage cost
// more than it saved).  The host plan, which only the runtime-length path
// follows, keeps them apart.
constexpr RadPlan make_radplan(int P)
{
    RadPlan rp{0, {0, 0, 0, 0, 0, 0, 0, 0}};
    int m = P, e2 = 0, e3 = 0, e5 = 0;
    while (m % 2 == 0) { m /= 2; e2++; }
    while (m % 3 == 0) { m /= 3; e3++; }
    while (m % 5 == 0) { m /= 5; e5++; }
    while (e2 >= 3) { rp.r[rp.n++] = 8; e2 -= 3; }
    if (e2 == 2) rp.r[rp.n++] = 4;
    if (e2 == 1 && e5 > 0 && P / 10 >= 32) { rp.r[rp.n++] = 10; e5--; }
    else if (e2 == 1) rp.r[rp.n++] = 2;
    for (int k = 0; k < e5; k++) rp.r[rp.n++] = 5;
    for (int k = 0; k < e3; k++) rp.r[rp.n++] = 3;
    return rp;
}

template <int P, int R, int NS>
__device__ __forceinline__ void fft_stage_t(double2* z, const double2* __restrict__ tw, int lane)
{
    constexpr int NB = P / R;
    constexpr int TMAX = (NB + WAVE - 1) / WAVE;
    constexpr int TWS = P / (NS * R);
    double2 v[TMAX][R];
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        const int j = lane + WAVE * t;
        if (NB % WAVE == 0 || j < NB) {
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = z[pz(j + r * NB)];
        }
    }
    wave_sync();
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        const int j = lane + WAVE * t;
        if (NB % WAVE == 0 || j < NB) {
            const int k = j % NS;
#pragma unroll
            for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], tw[k * r * TWS]);
            dft_r<R>(v[t]);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) z[pz(base + r * NS)] = v[t][r];
        }
    }
    wave_sync();
}

template <int P, int ST, int NS>
__device__ __forceinline__ void fft_stages_t(double2* z, const double2* __restrict__ tw, int lane)
{
    constexpr RadPlan rp = make_radplan(P);
    if constexpr (ST < rp.n) {
        constexpr int R = rp.r[ST];
        fft_stage_t<P, R, NS>(z, tw, lane);
        fft_stages_t<P, ST + 1, NS * R>(z, tw, lane);
    }
}

template <int P>
__device__ __forceinline__ double pre_y_t(const double* s, const double* sn, int j)
{
    constexpr int M = 2 * P;
    const int jm = M - j;
    const double xa = (j >= 1) ? s[j - 1] : 0.0;
    const double xb = (jm >= 1 && jm <= M - 1) ? s[jm - 1] : 0.0;
    const double sj = sn[j < jm ? j : jm];
    return (j == 0) ? 0.0 : fma(sj, xa + xb, 0.5 * (xa - xb));
}

template <int P>
__device__ __forceinline__ void dst1_line_t(double* s, const DstTab& tb, int lane)
{
    asm volatile("" : "+v"(lane));
    constexpr RadPlan rp = make_radplan(P);
    constexpr int R0 = rp.n > 0 ? rp.r[0] : 1;
    double2* z = reinterpret_cast<double2*>(s);
    // --- pre-processing fused into stage 1 (Ns = 1, no twiddles)
    {
        constexpr int NB = P / R0;
        constexpr int TMAX = (NB + WAVE - 1) / WAVE;
        double2 v[TMAX][R0];
#pragma unroll
        for (int t = 0; t < TMAX; t++) {
            const int j = lane + WAVE * t;
            if (NB % WAVE == 0 || j < NB) {
#pragma unroll
                for (int r = 0; r < R0; r++) {
                    const int m = j + r * NB;
                    v[t][r] = make_double2(pre_y_t<P>(s, tb.sn, 2 * m), pre_y_t<P>(s, tb.sn, 2 * m + 1));
                }
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < TMAX; t++) {
            const int j = lane + WAVE * t;
            if (NB % WAVE == 0 || j < NB) {
                dft_r<R0>(v[t]);
#pragma unroll
                for (int r = 0; r < R0; r++) z[pz(j * R0 + r)] = v[t][r];
            }
        }
        wave_sync();
    }
    fft_stages_t<P, 1, R0>(z, tb.tw, lane);
    // --- real-DFT unpacking into R'_k at s[k] and Y(2k) at s[P + k]
    {
        constexpr int NK = P / 2 + 1;
        constexpr int KMAX = (NK + WAVE - 1) / WAVE;
        double2 za[KMAX], zb[KMAX];
#pragma unroll
        for (int t = 0; t < KMAX; t++) {
            const int k = lane + WAVE * t;
            if (k < NK) {
                const int k2 = (P - k) % P;
                za[t] = z[pz(k)];
                zb[t] = z[pz(k2)];
            }
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < KMAX; t++) {
            const int k = lane + WAVE * t;
            if (k < NK) {
                const int k2 = (P - k) % P;
                const double2 a = za[t], b = zb[t];
                {
                    double2 E = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
                    double2 O = make_double2(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));
                    double2 F = cadd(E, cmul(tb.wm[k], O));
                    s[k] = (k == 0) ? F.x : 2.0 * F.x;
                    if (k > 0) s[P + k] = -2.0 * F.y;
                }
                if (k2 != k) {
                    double2 E = make_double2(0.5 * (b.x + a.x), 0.5 * (b.y - a.y));
                    double2 O = make_double2(0.5 * (b.y + a.y), -0.5 * (b.x - a.x));
                    double2 F = cadd(E, cmul(tb.wm[k2], O));
                    s[k2] = 2.0 * F.x;
                    s[P + k2] = -2.0 * F.y;
                }
            }
        }
        wave_sync();
    }
    // --- scan of R' in rounds of 64, natural-order output s[n] = Y(n+1)
    {
        constexpr int NR = (P + WAVE - 1) / WAVE;
        double sc[NR], ev[NR];
        double carry = 0.0;
#pragma unroll
        for (int t = 0; t < NR; t++) {
            const int k = WAVE * t + lane;
            const double v = (P % WAVE == 0 || k < P) ? s[k] : 0.0;
            const double incl = wave_incl_scan(v, lane);
            sc[t] = carry + incl;
            carry += wave_last(incl);
            ev[t] = (k >= 1 && (P % WAVE == 0 || k < P)) ? s[P + k] : 0.0;
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < NR; t++) {
            const int k = WAVE * t + lane;
            if (P % WAVE == 0 || k < P) {
                s[2 * k] = sc[t];
                if (k >= 1) s[2 * k - 1] = ev[t];
            }
        }
        wave_sync();
    }
}

// The transform lengths with a compile-time specialisation (P = M/2); any
// other length runs dst1_line.
__host__ __device__ inline bool dst_specialised(int P)
{
    return P == 512 || P == 400 || P == 300 || P == 256 || P == 200 || P == 150 || P == 128 || P == 100;
}

// ---------------------------------------------------------------------------
// DST-I of lines of 1023 values (M = 1024, P = 512 = 8^3): the transform of
// grids with 1024 cells along a direction.  Same algorithm as dst1_line, laid
// out so that nothing branches per lane:
//  * guarded line layout: element e = 0..M of the line x_e lives at gsl(e)
//    (two doubles of padding per 16), with guard cells x_0 = x_M = 0, so the
//    pre-processing reads x_j and x_{M-j} unconditionally;
//  * stage 1 takes its radix-8 inputs y_{2m}, y_{2m+1} (m = lane + 64 r) from
//    one 16-byte read of (x_{2m}, x_{2m+1}) plus the mirrored pair;
//  * the unpacking gives lane L the eight consecutive k = 8L..8L+7, so the
//    prefix sum of the odd outputs is a serial sum in the lane plus one wave
//    scan (instead of eight), and the lane owns outputs n = 16L..16L+15.
// On exit gsl(n) holds Y_n (n = 1..M-1) and the guards are 0 again.
// ---------------------------------------------------------------------------
namespace f1024 {
constexpr int M = 1024, P = 512;
__host__ __device__ constexpr int gsl(int e) { return e + 2 * (e >> 4); }
// doubles of one line slot: gsl(M) + 1 = 1153 for the real layout and
// 2 (pz(P - 1) + 1) = 1150 for the complex work array, rounded to even, + 2
// to stagger the LDS banks of neighbouring slots
constexpr int SLOT = 1156;
// doubles of the LDS tables: the stage twiddles (P complex: W_P^(8kr) of
// stage 2 at 8r + k, W_P^(jr) of stage 3 at 64r + j, so that the 64 lanes of
// a read hit consecutive entries), W_M^k at k + k/8 (P + P/8 complex),
// sin(pi j/M) for j <= M/2 (rounded to even), then the mirrored pairs
// {sin(pi (M-2m)/M), sin(pi (M-2m-1)/M)} of stage 1 for m = P/2 .. P-1 (P doubles)
constexpr int TWP = P + P / 8;
constexpr int SNM = M / 2 + 2;
constexpr int TAB = 2 * P + 2 * TWP + SNM + P;
}  // namespace f1024

__device__ inline DstTab load_dst_tab_1024(double* lds, const DstPlan& pl, int tid, int nthreads)
{
    using namespace f1024;
    double2* tw = reinterpret_cast<double2*>(lds);
    double2* wm = tw + P;
    double* sn = reinterpret_cast<double*>(wm + TWP);
    double* snm = sn + SNM;
    for (int i = tid; i < P; i += nthreads) {
        // i < 64: stage 2, r = i / 8, k = i % 8 -> W_P^(8 k r); else stage 3, r = i / 64, j = i % 64 -> W_P^(j r)
        const int e = i < 64 ? 8 * (i & 7) * (i >> 3) : (i & 63) * (i >> 6);
        tw[i] = pl.tw[e];
        wm[i + (i >> 3)] = pl.wm[i];
    }
    for (int i = tid; i <= M / 2; i += nthreads) sn[i] = pl.sn[i];
    for (int i = tid; i < P; i += nthreads) {  // m = P/2 + i/2: sin(pi (M - 2m - (i & 1))/M)
        const int m = P / 2 + (i >> 1);
        snm[i] = pl.sn[M - 2 * m - (i & 1)];
    }
    DstTab t;
    t.tw = tw;
    t.wm = wm;
    t.sn = sn;
    return t;
}

// The same tables in two phases, so that a thread has all its global loads
// in flight at once (a copy loop that stores each value as it arrives waits
// for every load in turn): tab1024_issue loads a thread's share into
// registers, tab1024_store writes it to LDS.  NT threads take part; other
// loads may be issued between the two calls.
template <int NT>
struct Tab1024Regs {
    static constexpr int A = (f1024::P + NT - 1) / NT;          // tw, wm, mirrored-sine entries
    static constexpr int B = (f1024::M / 2 + 1 + NT - 1) / NT;  // sine entries
    double2 tw[A], wm[A];
    double sm[A], sn[B];
};

template <int NT>
__device__ __forceinline__ void tab1024_issue(Tab1024Regs<NT>& R, const DstPlan& pl, int tid)
{
    using namespace f1024;
    // every register is loaded (indices past the end are clamped and their
    // values never stored), which keeps the arrays out of scratch memory
#pragma unroll
    for (int q = 0; q < Tab1024Regs<NT>::A; q++) {
        const int i = min(tid + NT * q, P - 1);
        const int e = i < 64 ? 8 * (i & 7) * (i >> 3) : (i & 63) * (i >> 6);
        const int m = P / 2 + (i >> 1);
        R.tw[q] = pl.tw[e];
        R.wm[q] = pl.wm[i];
        R.sm[q] = pl.sn[M - 2 * m - (i & 1)];
    }
#pragma unroll
    for (int q = 0; q < Tab1024Regs<NT>::B; q++) R.sn[q] = pl.sn[min(tid + NT * q, M / 2)];
}

template <int NT>
__device__ __forceinline__ DstTab tab1024_store(double* lds, const Tab1024Regs<NT>& R, int tid)
{
    using namespace f1024;
    double2* tw = reinterpret_cast<double2*>(lds);
    double2* wm = tw + P;
    double* sn = reinterpret_cast<double*>(wm + TWP);
    double* snm = sn + SNM;
#pragma unroll
    for (int q = 0; q < Tab1024Regs<NT>::A; q++) {
        const int i = tid + NT * q;
        if (i < P) {
            tw[i] = R.tw[q];
            wm[i + (i >> 3)] = R.wm[q];
            snm[i] = R.sm[q];
        }
    }
#pragma unroll
    for (int q = 0; q < Tab1024Regs<NT>::B; q++) {
        const int i = tid + NT * q;
        if (i <= M / 2) sn[i] = R.sn[q];
    }
    DstTab t;
    t.tw = tw;
    t.wm = wm;
    t.sn = sn;
    return t;
}

// exclusive prefix sum over the lanes (lane 0 gets 0): wave_shr:1, then scan
__device__ __forceinline__ double wave_excl_scan(double v)
{
    return wave_incl_scan(dpp_d<0x138, 0xf>(v), 0);
}

// `out(L, i, ye, yo)` stores the outputs n = 16L + 2i (ye) and n + 1 (yo) of
// lane L; it may scale them first (ye of n = 0 is the guard and must stay 0).
template <class OUT>
__device__ __forceinline__ void dst1024_line(double* S, const DstTab& tb, int lane, OUT out)
{
    using namespace f1024;
    asm volatile("" : "+v"(lane));
    double2* z = reinterpret_cast<double2*>(S);
    const int j = lane;
    double2 v[8];
    // --- stage 1 (Ns = 1): pre-processing + radix 8, inputs m = j + 64 r
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const int m = j + 64 * r;
        const double2 a = *reinterpret_cast<const double2*>(S + gsl(2 * m));  // x_{2m}, x_{2m+1}
        const double b0 = S[gsl(M - 2 * m)];                                    // x_{M-2m}
        const double b1 = S[gsl(M - 2 * m - 1)];                                // x_{M-2m-1}
        double s0, s1;
        if (r < 4) {  // 2m + 1 <= M/2: sin(pi (2m)/M), sin(pi (2m+1)/M) straight from the table
            const double2 t = *reinterpret_cast<const double2*>(tb.sn + 2 * m);
            s0 = t.x;
            s1 = t.y;
        } else {  // mirrored, sin(pi e/M) = sin(pi (M-e)/M): the pair stored for m
            const double2 t = *reinterpret_cast<const double2*>(tb.sn + SNM + 2 * (m - P / 2));
            s0 = t.x;
            s1 = t.y;
        }
        v[r] = make_double2(fma(s0, a.x + b0, 0.5 * (a.x - b0)), fma(s1, a.y + b1, 0.5 * (a.y - b1)));
    }
    wave_sync();
    dft8(v);
#pragma unroll
    for (int r = 0; r < 8; r++) z[9 * j + r] = v[r];  // pz(8 j + r)
    wave_sync();
    const int p0 = j + (j >> 3);  // pz(j); pz(j + 64 r) = p0 + 72 r
    // --- stage 2 (Ns = 8): twiddles W_P^(8 k r), k = j mod 8
    {
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = z[p0 + 72 * r];
        const int k = j & 7;
#pragma unroll
        for (int r = 1; r < 8; r++) v[r] = cmul(v[r], tb.tw[8 * r + k]);  // W_P^(8 k r)
        wave_sync();
        dft8(v);
        const int q0 = ((j >> 3) << 6) + k + ((j >> 3) << 3);  // pz(64 (j/8) + k); + 9 r per r
#pragma unroll
        for (int r = 0; r < 8; r++) z[q0 + 9 * r] = v[r];
        wave_sync();
    }
    // --- stage 3 (Ns = 64): twiddles W_P^(j r), natural-order output
    {
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = z[p0 + 72 * r];
#pragma unroll
        for (int r = 1; r < 8; r++) v[r] = cmul(v[r], tb.tw[64 * r + j]);  // W_P^(j r)
        wave_sync();
        dft8(v);
#pragma unroll
        for (int r = 0; r < 8; r++) z[p0 + 72 * r] = v[r];
        wave_sync();
    }
    // --- unpacking, lane L = j owns k = 8L + i: with a = z_k, b = z_{P-k},
    //     2 F_k = (a + conj b) + W_M^k (-i)(a - conj b) (exactly twice the F_k
    //     of dst1_line); R'_k = 2 R_k (R'_0 = R_0) and Y(2k) = 2 I_k = -2 Im F_k
    double rr[8], ye[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = 8 * j + i;
        const double2 a = z[9 * j + i];
        const int k2 = (P - k) & (P - 1);
        const double2 b = z[k2 + (k2 >> 3)];
        const double2 w = tb.wm[9 * j + i];
        const double2 E2 = make_double2(a.x + b.x, a.y - b.y);
        const double2 O2 = make_double2(a.y + b.y, -(a.x - b.x));
        const double2 F2 = cadd(E2, cmul(w, O2));
        rr[i] = F2.x;
        ye[i] = -F2.y;
    }
    wave_sync();
    if (j == 0) {
        rr[0] = 0.5 * rr[0];  // R'_0 = R_0
        ye[0] = 0.0;          // n = 0: the guard cell x_0
    }
    // --- Y(2k+1) = sum_{m <= k} R'_m: serial in the lane, one wave scan across
    double cs[8];
    cs[0] = rr[0];
#pragma unroll
    for (int i = 1; i < 8; i++) cs[i] = cs[i - 1] + rr[i];
    const double off = wave_excl_scan(cs[7]);
#pragma unroll
    for (int i = 0; i < 8; i++) out(j, i, ye[i], off + cs[i]);
    wave_sync();
}

// plain store of a dst1024_line output pair into the guarded slot
struct F1024Store {
    double* S;
    __device__ __forceinline__ void operator()(int L, int i, double ye, double yo) const
    {
        *reinterpret_cast<double2*>(S + f1024::gsl(16 * L + 2 * i)) = make_double2(ye, yo);
    }
};

// ---------------------------------------------------------------------------
// The 1023-point DST-I split over two wavefronts that run at the same time,
// for launches that hold one line per SIMD (a single level): each wavefront
// transforms half of the outputs with a 256-point FFT, so a line's transform
// is two independent chains of half the length.  Pairing x_j with x_{M-j}
// (M = 1024, P = 512):
//  * even outputs: Y_{2q} = 2 sum_{j<P} u_j sin(pi j q / P), u_j = x_j - x_{M-j},
//    the 511-point DST-I of u, by the half-length algorithm of dst1024_line
//    at P' = 256 (pre-processing, FFT, unpacking, prefix sum);
//  * odd outputs: Y_{2q+1} = (-1)^q C_q with C the DCT-III (REDFT01) of
//    X'_0 = 2 x_P, X'_t = x_{P-t} + x_{P+t} (t < P), by the half-length
//    real-IDFT form of Makhoul's algorithm: H_t = (X'_t - i X'_{P-t}) e^{i pi t/M}
//    (X'_P = x_0 + x_M = 0), Z_k = (H_k + H_{k+256}) + i e^{2 pi i k/P} (H_k - H_{k+256}),
//    z = IDFT_256(Z) (unnormalised, as conj(FFT(conj Z))), w_{2m} = Re z_m,
//    w_{2m+1} = Im z_m, C_{2r} = w_r, C_{2r+1} = w_{P-1-r}.  No prefix sum.
// The 256-point FFTs are four radix-4 Stockham stages, one butterfly per
// lane.  Lane L of either wavefront ends with the outputs n = 16 L + nl of its
// parity, which it hands to out(L, nl, value).  The transform reads the
// guarded line S (gsl layout, guards zero), works in its own padded complex
// array z (fsplit::ZS doubles) and writes nothing else: the caller places the
// outputs (out) where no wavefront of the line still reads.
// ---------------------------------------------------------------------------
namespace fsplit {
constexpr int H = 256;  // FFT length
// LDS tables, double2 units: W_512^k (k < 256) at WM + pw(k) (one pad per 4,
// so that lanes reading 4 consecutive k each hit distinct banks); the radix-4
// stage twiddles W_{4Ns}^{k r} (r = 1..3, k < Ns) at TW + off(Ns) + (r-1) Ns + k
// for Ns = 4, 16, 64 (off 0, 12, 60); the even half's pre-processing sines
// {sin(pi 2m/P), sin(pi (2m+1)/P)} at SA + m; then sin(pi j/M), j <= P, as doubles.
__host__ __device__ constexpr int pw(int k) { return k + (k >> 2); }
constexpr int WM = 0, TW = pw(H), SA = TW + 252, SN2 = SA + H;
constexpr int TAB = 2 * SN2 + 514;  // doubles
// doubles of one wavefront's complex work array: pz(H - 1) + 1 = 287 complex, + 2
constexpr int ZS = 576;
}  // namespace fsplit

struct SplitTab {
    const double2* wm;
    const double2* tw;
    const double2* sa;
    const double* sn;
};

// The split tables from the P = 512 plan (tw = W_512^e, sn = sin(pi j/1024)),
// in two phases as tab1024_issue / tab1024_store: NT threads take part.
template <int NT>
struct SplitTabRegs {
    static constexpr int A = (fsplit::H + NT - 1) / NT;        // W_512^k, pre-processing sines
    static constexpr int B = (252 + NT - 1) / NT;              // stage twiddles
    static constexpr int C = (f1024::M / 2 + 1 + NT - 1) / NT;  // sines
    double2 wm[A], sa[A], tw[B];
    double sn[C];
};

template <int NT>
__device__ __forceinline__ void splittab_issue(SplitTabRegs<NT>& R, const DstPlan& pl, int tid)
{
    // indices past the end are clamped (their values are never stored)
#pragma unroll
    for (int q = 0; q < SplitTabRegs<NT>::A; q++) {
        const int m = min(tid + NT * q, fsplit::H - 1);
        R.wm[q] = pl.tw[m];
        const int e0 = 4 * m, e1 = 4 * m + 2;  // sin(pi i / 512) = sin(pi 2i / 1024), mirrored past 512
        R.sa[q] = make_double2(pl.sn[min(e0, f1024::M - e0)], pl.sn[min(e1, f1024::M - e1)]);
    }
#pragma unroll
    for (int q = 0; q < SplitTabRegs<NT>::B; q++) {
        const int i = min(tid + NT * q, 251);
        // i < 12: Ns = 4; i < 60: Ns = 16; else Ns = 64.  W_{4Ns}^{k r} = W_512^{k r 128 / Ns}
        const int ns = i < 12 ? 4 : (i < 60 ? 16 : 64);
        const int o = i < 12 ? i : (i < 60 ? i - 12 : i - 60);
        const int r = o / ns + 1, k = o % ns;
        R.tw[q] = pl.tw[k * r * (128 / ns)];
    }
#pragma unroll
    for (int q = 0; q < SplitTabRegs<NT>::C; q++) R.sn[q] = pl.sn[min(tid + NT * q, f1024::M / 2)];
}

template <int NT>
__device__ __forceinline__ SplitTab splittab_store(double* lds, const SplitTabRegs<NT>& R, int tid)
{
    using namespace fsplit;
    double2* t2 = reinterpret_cast<double2*>(lds);
    double* sn = lds + 2 * SN2;
#pragma unroll
    for (int q = 0; q < SplitTabRegs<NT>::A; q++) {
        const int m = tid + NT * q;
        if (m < H) {
            t2[WM + pw(m)] = R.wm[q];
            t2[SA + m] = R.sa[q];
        }
    }
#pragma unroll
    for (int q = 0; q < SplitTabRegs<NT>::B; q++) {
        const int i = tid + NT * q;
        if (i < 252) t2[TW + i] = R.tw[q];
    }
#pragma unroll
    for (int q = 0; q < SplitTabRegs<NT>::C; q++) {
        const int i = tid + NT * q;
        if (i <= f1024::M / 2) sn[i] = R.sn[q];
    }
    SplitTab t;
    t.wm = t2 + WM;
    t.tw = t2 + TW;
    t.sa = t2 + SA;
    t.sn = sn;
    return t;
}

// radix-4 Stockham stage Ns of the 256-point FFT, butterfly j = lane.  The
// reads pz(j + 64 r) = pz(j) + 72 r; the writes pz(base + r Ns) are a lane
// base plus a constant: Ns = 4: base + 4r + r/2, Ns = 16: base + 18 r,
// Ns = 64: base + 72 r (base + r Ns stays inside one 8-group for r Ns < 8
// and crosses whole groups otherwise).
template <int NS>
__device__ __forceinline__ void fft256_stage(double2* z, const double2* __restrict__ tw, int j, int pj)
{
    double2 v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = z[pj + 72 * r];
    const int k = j & (NS - 1);
#pragma unroll
    for (int r = 1; r < 4; r++) v[r] = cmul(v[r], tw[(r - 1) * NS + k]);
    wave_sync();
    dft4(v);
    const int base = (j - k) * 4 + k;  // (j / Ns) Ns 4 + k
    if constexpr (NS == 4) {
        const int pb = pz(base);  // base = 16 (j/4) + k, k < 4: pz(base + 4r) = pz(base) + 4r + r/2
#pragma unroll
        for (int r = 0; r < 4; r++) z[pb + 4 * r + (r >> 1)] = v[r];
    } else if constexpr (NS == 16) {
        const int pb = pz(base);  // base = 64 (j/16) + k, k < 16: pz(base + 16r) = pz(base) + 18r
#pragma unroll
        for (int r = 0; r < 4; r++) z[pb + 18 * r] = v[r];
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++) z[pj + 72 * r] = v[r];  // base = j
    }
    wave_sync();
}

// half 0: even outputs, half 1: odd outputs (wave-uniform).  Line reads at
// m = j + 64 r (or 2m) are a lane base plus a constant: gsl(e +- 64 s) =
// gsl(e) +- 72 s for e, e +- 64 s >= 0.
template <class OUT>
__device__ __forceinline__ void dst1024_split(const double* S, double2* z, const SplitTab& tb, int lane, int half, OUT out)
{
    using f1024::gsl;
    asm volatile("" : "+v"(lane));
    half = __builtin_amdgcn_readfirstlane(half);  // uniform branches below
    const int j = lane;
    double2 v[4];
    // --- stage 1 (Ns = 1) inputs m = j + 64 r, formed from the line
    if (half == 0) {
        // z_m = y_{2m} + i y_{2m+1}, y_i = sin(pi i/P) (u_i + u_{P-i}) + (u_i - u_{P-i}) / 2
        const double* pa = S + gsl(2 * j);          // x_{2m}, x_{2m+1}: + 144 r
        const double* pc = S + gsl(512 + 2 * j);    // x_{P+2m}, x_{P+2m+1}
        const double* pb0 = S + gsl(1024 - 2 * j);  // x_{M-2m}: - 144 r
        const double* pb1 = S + gsl(1023 - 2 * j);  // x_{M-2m-1}
        const double* pd0 = S + gsl(512 - 2 * j);   // x_{P-2m}
        const double* pd1 = S + gsl(511 - 2 * j);   // x_{P-2m-1}
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double2 xa = *reinterpret_cast<const double2*>(pa + 144 * r);
            const double2 xc = *reinterpret_cast<const double2*>(pc + 144 * r);
            const double xb0 = pb0[-144 * r], xb1 = pb1[-144 * r];
            const double xd0 = pd0[-144 * r], xd1 = pd1[-144 * r];
            const double2 sv = tb.sa[j + 64 * r];
            const double u0 = xa.x - xb0, w0 = xd0 - xc.x;  // u_{2m}, u_{P-2m}
            const double u1 = xa.y - xb1, w1 = xd1 - xc.y;  // u_{2m+1}, u_{P-2m-1}
            v[r] = make_double2(fma(sv.x, u0 + w0, 0.5 * (u0 - w0)), fma(sv.y, u1 + w1, 0.5 * (u1 - w1)));
        }
    } else {
        // conj(Z_m)
        const double* q0 = S + gsl(512 - j);   // x_{P-m}: - 72 r
        const double* q1 = S + gsl(512 + j);   // x_{P+m}: + 72 r
        const double* q2 = S + gsl(j);         // x_m
        const double* q3 = S + gsl(1024 - j);  // x_{M-m}
        const double* q4 = S + gsl(256 - j);   // x_{256-m}
        const double* q5 = S + gsl(768 + j);   // x_{768+m}
        const double* q6 = S + gsl(256 + j);   // x_{256+m}
        const double* q7 = S + gsl(768 - j);   // x_{768-m}
        const double* sn = tb.sn;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int d72 = 72 * r, d64 = 64 * r;
            const double a0 = q0[-d72] + q1[d72];  // X'_m
            const double b0 = q2[d72] + q3[-d72];  // X'_{P-m}
            const double a1 = q4[-d72] + q5[d72];  // X'_{m+256}
            const double b1 = q6[d72] + q7[-d72];  // X'_{256-m}
            const double c0 = sn[512 - j - d64], s0 = sn[j + d64];        // cos, sin(pi m/M)
            const double c1 = sn[256 - j - d64], s1 = sn[256 + j + d64];  // cos, sin(pi (m+256)/M)
            const double2 h0 = make_double2(fma(a0, c0, b0 * s0), fma(a0, s0, -(b0 * c0)));
            const double2 h1 = make_double2(fma(a1, c1, b1 * s1), fma(a1, s1, -(b1 * c1)));
            const double2 d = csub(h0, h1);
            const double2 e = tb.wm[fsplit::pw(j) + 80 * r];  // W_P^m; i conj(W_P^m) d:
            const double2 t = make_double2(-fma(e.x, d.y, -(e.y * d.x)), fma(e.x, d.x, e.y * d.y));
            v[r] = make_double2((h0.x + h1.x) + t.x, -((h0.y + h1.y) + t.y));
        }
    }
    dft4(v);
    {
        const int p4 = 4 * j + (j >> 1);  // pz(4j + r) = p4 + r, r < 4
#pragma unroll
        for (int r = 0; r < 4; r++) z[p4 + r] = v[r];
    }
    wave_sync();
    const int pj = pz(j);
    fft256_stage<4>(z, tb.tw, j, pj);
    fft256_stage<16>(z, tb.tw + 12, j, pj);
    fft256_stage<64>(z, tb.tw + 60, j, pj);
    if (half == 0) {
        // unpacking, lane j owns k = 4j + i: 2 F_k = (a + conj b) + W_P^k (-i)(a - conj b),
        // a = z_k, b = z_{256-k}; the outputs 4k (2 I_k) and 4k + 2 (prefix sum of R')
        const int p4 = 4 * j + (j >> 1);  // pz(4j + i)
        // pz(256 - 4j - i): e0 = 256 - 4j, pz(e0) - i, one less for i >= 1 when e0 = 0 mod 8
        // (j even); lane 0 at i = 0 reads z_0
        const int e0 = 256 - 4 * j;
        const int pb0 = j == 0 ? 0 : e0 + (e0 >> 3);
        const int pb1 = e0 + (e0 >> 3) - ((j & 1) ? 0 : 1);
        double rr[4], ye[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double2 a = z[p4 + i];
            const double2 b = z[i == 0 ? pb0 : pb1 - i];
            const double2 w = tb.wm[5 * j + i];  // pw(4j + i)
            const double2 E2 = make_double2(a.x + b.x, a.y - b.y);
            const double2 O2 = make_double2(a.y + b.y, -(a.x - b.x));
            const double2 F2 = cadd(E2, cmul(w, O2));
            rr[i] = F2.x;
            ye[i] = -F2.y;
        }
        if (j == 0) {
            rr[0] = 0.5 * rr[0];  // R'_0 = R_0
            ye[0] = 0.0;          // n = 0: the guard cell x_0
        }
        double cs[4];
        cs[0] = rr[0];
#pragma unroll
        for (int i = 1; i < 4; i++) cs[i] = cs[i - 1] + rr[i];
        const double off = wave_excl_scan(cs[3]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            out(j, 4 * i, ye[i]);
            out(j, 4 * i + 2, off + cs[i]);
        }
    } else {
        // F = FFT(conj Z), z = conj F: C_q of q = 8j + t from F_{2j}, F_{2j+1},
        // F_{254-2j}, F_{255-2j}; output n = 2q + 1 = 16 j + 2t + 1 is (-1)^t C_q
        const int pa = 2 * j + (j >> 2);                    // pz(2j) (pz(2j + 1) = pa + 1)
        const int pc = (254 - 2 * j) + ((255 - 2 * j) >> 3);  // pz(254 - 2j) (pz(255 - 2j) = pc + 1)
        const double2 a = z[pa], b = z[pa + 1];
        const double2 c = z[pc], d = z[pc + 1];
        out(j, 1, a.x);
        out(j, 3, d.y);
        out(j, 5, -a.y);
        out(j, 7, -d.x);
        out(j, 9, b.x);
        out(j, 11, c.y);
        out(j, 13, -b.y);
        out(j, 15, -c.x);
    }
}

// ---------------------------------------------------------------------------
// get_bc! evaluated pointwise (src/fluid-operators/mg-utils.jl:7-32).
// r: coarse-level (nx-1) x (ny-1) array.  k: 1-based index into the boundary
// vector (offsets LEFT=0, RIGHT=ny+1, BOT=2(ny+1), TOP=2(ny+1)+nx+1).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rat(const double* r, const GridDev& g, int i, int j)  // r[i,j], 1-based
{
    return r[(i - 1) + (long)g.n1 * (j - 1)];
}

// One side of get_bc!: sample idx (1-based) of the coarse level along i at
// row jr (BOT/TOP) or along j at column ir (LEFT/RIGHT).  Odd idx take the
// coarse value at offset idx/2 from nx/4 (ny/4); even idx the mean of it and
// its lower neighbour.  Both values are loaded and selected, so no load sits
// in a lane-divergent branch.
__device__ __forceinline__ double bc_row(const double* __restrict__ r, const GridDev& g, int jr, int idx)
{
    const double* p = r + (g.nx / 4 + idx / 2 - 1) + (long)g.n1 * (jr - 1);
    const double a = p[0], b = p[-1];
    return (idx & 1) ? a : 0.5 * (a + b);
}

__device__ __forceinline__ double bc_col(const double* __restrict__ r, const GridDev& g, int ir, int idx)
{
    const double* p = r + (ir - 1) + (long)g.n1 * (g.ny / 4 + idx / 2 - 1);
    const double a = p[0], b = p[-(long)g.n1];
    return (idx & 1) ? a : 0.5 * (a + b);
}

__device__ __forceinline__ double get_bc_val(const double* __restrict__ r, const GridDev& g, int k)
{
    if (k <= g.RIGHT) return bc_col(r, g, g.nx / 4, k - g.LEFT);
    if (k <= g.BOT) return bc_col(r, g, 3 * g.nx / 4, k - g.RIGHT);
    if (k <= g.TOP) return bc_row(r, g, g.ny / 4, k - g.BOT);
    return bc_row(r, g, 3 * g.ny / 4, k - g.TOP);
}

// get_bc_val without branches, for code that evaluates many boundary values
// at once: the coarse sample address of boundary index k is selected, both
// samples are loaded, the value is get_bc_val's expression.  With !valid the
// loads go to r0 (any field of the grid's size) and the value is zero.
__device__ __forceinline__ double get_bc_val_sel(const double* __restrict__ r, const double* __restrict__ r0,
                                                 bool valid, const GridDev& g, int k)
{
    const bool col = k <= g.BOT;                            // LEFT / RIGHT: a column of the coarse level
    const bool first = col ? (k <= g.RIGHT) : (k <= g.TOP);  // LEFT / BOT
    const int idx = k - (col ? (first ? g.LEFT : g.RIGHT) : (first ? g.BOT : g.TOP));
    const int ir = first ? g.nx / 4 : 3 * g.nx / 4;
    const int jr = first ? g.ny / 4 : 3 * g.ny / 4;
    const long off = col ? (ir - 1) + (long)g.n1 * (g.ny / 4 + idx / 2 - 1) : (g.nx / 4 + idx / 2 - 1) + (long)g.n1 * (jr - 1);
    const long back = col ? (long)g.n1 : 1;
    const double* p = (valid ? r : r0) + off;
    const double a = p[0], b = p[-back];
    return valid ? ((idx & 1) ? a : 0.5 * (a + b)) : 0.0;
}

// ---------------------------------------------------------------------------
// curl! on the multigrid (src/fluid-operators/curl.jl:81-116) at one flux of
// one level: psi is the level's streamfunction, its boundary values
// get_bc!(pc) of the next coarser level's psi (zero when !hasbc, the
// coarsest level).  Every kernel that needs a flux from psi calls these, so
// the value is the same bit for bit wherever it is formed.
// ---------------------------------------------------------------------------
// x-flux (i, j), i = 1..nx+1, j = 1..ny
__device__ __forceinline__ double curl_qx(const double* __restrict__ psi, const double* __restrict__ pc, bool hasbc,
                                          const GridDev& g, int i, int j)
{
    auto PS = [&](int ii, int jj) -> double { return psi[(ii - 1) + (long)g.n1 * (jj - 1)]; };
    auto BC = [&](int k) -> double { return hasbc ? get_bc_val(pc, g, k) : 0.0; };
    if (i >= 2 && i <= g.nx && j >= 2 && j <= g.ny - 1) return PS(i - 1, j) - PS(i - 1, j - 1);
    if (i == 1) return BC((g.LEFT + 1) + j) - BC(g.LEFT + j);
    if (i == g.nx + 1) return BC((g.RIGHT + 1) + j) - BC(g.RIGHT + j);
    if (j == 1) return PS(i - 1, 1) - BC((g.BOT - 1) + i);
    return BC(i + g.TOP) - PS(i - 1, g.ny - 1);  // j == ny
}

// curl_qx and curl_qy without branches: the two operands of the difference
// (a ψ value or a boundary value each) are both loaded and selected, so that
// a thread evaluating many fluxes has all their loads in flight at once.
// The value is the branching form's, bit for bit.
__device__ __forceinline__ double curl_qx_sel(const double* __restrict__ psi, const double* __restrict__ pc, bool hasbc,
                                              const GridDev& g, int i, int j)
{
    const int nx = g.nx, ny = g.ny;
    // operand a: ψ(i-1, j) [interior, j = 1 -> ψ(i-1, 1)] or bc; operand b: ψ(i-1, j-1) [interior,
    // j = ny -> ψ(i-1, ny-1) as a] or bc
    const bool in = i >= 2 && i <= nx;
    const bool bl = i == 1, br = i == nx + 1;
    // boundary indices of the operands that are boundary values
    const int ka = bl ? g.LEFT + 1 + j : br ? g.RIGHT + 1 + j : i + g.TOP;            // a is bc: i edge, or j == ny
    const int kb = bl ? g.LEFT + j : br ? g.RIGHT + j : (g.BOT - 1) + i;             // b is bc: i edge, or j == 1
    const bool abc = !in || j == ny;
    const bool bbc = !in || j == 1;
    // ψ operands (clamped into the field; used only when interior)
    const int ic = min(max(i - 1, 1), nx - 1);
    const int ja = j == ny ? ny - 1 : min(max(j, 1), ny - 1);   // j == ny: the ψ operand is b's, ψ(i-1, ny-1)
    const int jb = min(max(j - 1, 1), ny - 1);
    const double pa = psi[(ic - 1) + (long)g.n1 * (ja - 1)];
    const double pb = psi[(ic - 1) + (long)g.n1 * (jb - 1)];
    const double ca = get_bc_val_sel(pc, psi, hasbc, g, abc ? ka : g.LEFT + 1);
    const double cb = get_bc_val_sel(pc, psi, hasbc, g, bbc ? kb : g.LEFT + 1);
    // j == ny (interior i): a = bc(i + TOP), b = ψ(i-1, ny-1) (loaded as pa)
    if (in && j == ny) return ca - pa;
    return (abc ? ca : pa) - (bbc ? cb : pb);
}

__device__ __forceinline__ double curl_qy_sel(const double* __restrict__ psi, const double* __restrict__ pc, bool hasbc,
                                              const GridDev& g, int i, int j)
{
    const int nx = g.nx, ny = g.ny;
    const bool jn = j >= 2 && j <= ny;
    const bool bb = j == 1, bt = j == ny + 1;
    // j edge: bc - bc; i == 1: bc(LEFT + j) - ψ(1, j-1); i == nx: ψ(nx-1, j-1) - bc(RIGHT + j)
    const int ka = bb ? g.BOT + i : bt ? g.TOP + i : g.LEFT + j;
    const int kb = bb ? (g.BOT + 1) + i : bt ? (g.TOP + 1) + i : g.RIGHT + j;
    const bool abc = !jn || i == 1;
    const bool bbc = !jn || i == nx;
    const int jc = min(max(j - 1, 1), ny - 1);
    const int ia = i == nx ? nx - 1 : min(max(i - 1, 1), nx - 1);  // i == nx: ψ(nx-1, j-1) is operand a
    const int ib = min(max(i, 1), nx - 1);                          // i == 1: ψ(1, j-1) is operand b
    const double pa = psi[(ia - 1) + (long)g.n1 * (jc - 1)];
    const double pb = psi[(ib - 1) + (long)g.n1 * (jc - 1)];
    const double ca = get_bc_val_sel(pc, psi, hasbc, g, abc ? ka : g.LEFT + 1);
    const double cb = get_bc_val_sel(pc, psi, hasbc, g, bbc ? kb : g.LEFT + 1);
    return (abc ? ca : pa) - (bbc ? cb : pb);
}

// y-flux (i, j), i = 1..nx, j = 1..ny+1
__device__ __forceinline__ double curl_qy(const double* __restrict__ psi, const double* __restrict__ pc, bool hasbc,
                                          const GridDev& g, int i, int j)
{
    auto PS = [&](int ii, int jj) -> double { return psi[(ii - 1) + (long)g.n1 * (jj - 1)]; };
    auto BC = [&](int k) -> double { return hasbc ? get_bc_val(pc, g, k) : 0.0; };
    if (j >= 2 && j <= g.ny && i >= 2 && i <= g.nx - 1) return PS(i - 1, j - 1) - PS(i, j - 1);
    if (j == 1) return BC(g.BOT + i) - BC((g.BOT + 1) + i);
    if (j == g.ny + 1) return BC(g.TOP + i) - BC((g.TOP + 1) + i);
    if (i == 1) return BC(g.LEFT + j) - PS(1, j - 1);
    return PS(g.nx - 1, j - 1) - BC(g.RIGHT + j);  // i == nx
}

}  // namespace ibpm
