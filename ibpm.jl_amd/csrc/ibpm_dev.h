// ibpm_dev.h -- device-side building blocks shared by the HIP kernels.
//
// Layouts follow the reference exactly (Julia column-major, 1-based in the
// comments): a circulation-like field of one grid level is an (nx-1) x (ny-1)
// array; a flux field of one level is the x-fluxes as an (nx+1) x ny array
// followed by the y-fluxes as an nx x (ny+1) array
// (src/fluid-domain/fluid-domain.jl:192-202).  Levels are consecutive columns
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
    int radix[16];
    const double2* tw;  // exp(-2 pi i e / P), e < P
    const double* sn;   // kind 0: sin(pi j / M), j <= M/2;  kind 1: sin(pi t / M), t < 2M
    const double2* wm;  // exp(-2 pi i k / M), k < P
};

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

// One in-place Stockham stage of radix R over P complex points in LDS `z`,
// executed by one wavefront.  Ns = product of the radices of earlier stages.
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
            for (int r = 0; r < R; r++) v[t][r] = z[j + r * nb];
        }
    }
    wave_sync();
#pragma unroll
    for (int t = 0; t < TMAX; t++) {
        int j = lane + WAVE * t;
        if (j < nb) {
            int k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], tw[k * r * tws]);
            }
            if constexpr (R == 2) dft2(v[t]);
            else if constexpr (R == 3) dft3(v[t]);
            else if constexpr (R == 4) dft4(v[t]);
            else if constexpr (R == 5) dft5(v[t]);
            else dft8(v[t]);
            int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) z[base + r * Ns] = v[t][r];
        }
    }
    wave_sync();
}

// Wave-level inclusive prefix sum of one double per lane.
__device__ __forceinline__ double wave_incl_scan(double v, int lane)
{
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        double o = __shfl_up(v, off, WAVE);
        if (lane >= off) v += o;
    }
    return v;
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

__device__ __forceinline__ void dst1_line(double* s, const DstPlan& pl, int lane)
{
    const int M = pl.M;
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
                    a = fma(s[j - 1], pl.sn[e], a);
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
    const int P = pl.P;
    // --- pre-processing: pairs (j, M-j), j = 1..M/2, into y_j stored at s[j]
    {
        double xa[MAX_PAIRS], xb[MAX_PAIRS];
#pragma unroll
        for (int t = 0; t < MAX_PAIRS; t++) {
            int j = lane + 1 + WAVE * t;
            if (j <= P) {
                xa[t] = s[j - 1];
                xb[t] = (j < P) ? s[M - j - 1] : 0.0;
            }
        }
        wave_sync();
        if (lane == 0) s[0] = 0.0;  // y_0 = 0  (s[0] already consumed above)
#pragma unroll
        for (int t = 0; t < MAX_PAIRS; t++) {
            int j = lane + 1 + WAVE * t;
            if (j < P) {
                double sj = pl.sn[j];
                double a = xa[t] + xb[t];
                double d = 0.5 * (xa[t] - xb[t]);
                s[j] = fma(sj, a, d);
                s[M - j] = fma(sj, a, -d);
            } else if (j == P) {
                s[P] = 2.0 * xa[t];  // sin(pi/2) (x + x)
            }
        }
        wave_sync();
    }
    // --- complex FFT of P points in place (Stockham, natural order out)
    double2* z = reinterpret_cast<double2*>(s);
    int Ns = 1;
    for (int st = 0; st < pl.nst; st++) {
        int R = pl.radix[st];
        if (R == 8) fft_stage<8>(z, P, Ns, pl.tw, lane);
        else if (R == 4) fft_stage<4>(z, P, Ns, pl.tw, lane);
        else if (R == 2) fft_stage<2>(z, P, Ns, pl.tw, lane);
        else if (R == 5) fft_stage<5>(z, P, Ns, pl.tw, lane);
        else fft_stage<3>(z, P, Ns, pl.tw, lane);
        Ns *= R;
    }
    // --- real-DFT unpacking: for pairs (k, P-k), k = 0..P/2
    {
        constexpr int KMAX = (MAXP / 2 + 1 + WAVE - 1) / WAVE;
        double2 za[KMAX], zb[KMAX];
        const int nk = P / 2 + 1;
#pragma unroll
        for (int t = 0; t < KMAX; t++) {
            int k = lane + WAVE * t;
            if (k < nk) {
                int k2 = (P - k) % P;
                za[t] = z[k];
                zb[t] = z[k2];
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
                    double2 F = cadd(E, cmul(pl.wm[k], O));
                    s[2 * k] = (k == 0) ? F.x : 2.0 * F.x;      // R'_k at the slot of Y(2k+1)
                    if (k > 0) s[2 * k - 1] = -2.0 * F.y;       // Y(2k) = 2 I_k
                }
                if (k2 != k) {
                    double2 E = make_double2(0.5 * (b.x + a.x), 0.5 * (b.y - a.y));
                    double2 O = make_double2(0.5 * (b.y + a.y), -0.5 * (b.x - a.x));
                    double2 F = cadd(E, cmul(pl.wm[k2], O));
                    s[2 * k2] = 2.0 * F.x;
                    s[2 * k2 - 1] = -2.0 * F.y;
                }
            }
        }
        wave_sync();
    }
    // --- prefix sum over s[0], s[2], ..., s[2(P-1)]  ->  Y(1), Y(3), ...
    {
        const int C = (P + WAVE - 1) / WAVE;
        const int k0 = lane * C;
        const int k1 = min(k0 + C, P);
        double part = 0.0;
        for (int k = k0; k < k1; k++) part += s[2 * k];
        double incl = wave_incl_scan(part, lane);
        double run = incl - part;
        for (int k = k0; k < k1; k++) {
            run += s[2 * k];
            s[2 * k] = run;
        }
        wave_sync();
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

__device__ __forceinline__ double get_bc_val(const double* __restrict__ r, const GridDev& g, int k)
{
    const int nx = g.nx, ny = g.ny;
    int side, idx;  // idx = 1-based position on the side
    if (k <= g.RIGHT) { side = 0; idx = k - g.LEFT; }
    else if (k <= g.BOT) { side = 1; idx = k - g.RIGHT; }
    else if (k <= g.TOP) { side = 2; idx = k - g.BOT; }
    else { side = 3; idx = k - g.TOP; }
    if (side >= 2) {  // BOT / TOP : along i
        int jr = (side == 2) ? ny / 4 : 3 * ny / 4;
        int m = idx / 2;
        if (idx & 1) return rat(r, g, nx / 4 + m, jr);
        return 0.5 * (rat(r, g, nx / 4 + m, jr) + rat(r, g, nx / 4 + m - 1, jr));
    } else {  // LEFT / RIGHT : along j
        int ir = (side == 0) ? nx / 4 : 3 * nx / 4;
        int m = idx / 2;
        if (idx & 1) return rat(r, g, ir, ny / 4 + m);
        return 0.5 * (rat(r, g, ir, ny / 4 + m) + rat(r, g, ir, ny / 4 + m - 1));
    }
}

}  // namespace ibpm
