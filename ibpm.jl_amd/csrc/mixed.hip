// mixed.hip -- the mixed-precision form of the Laplacian inverses: every
// DST-I direction of Δinv and Ainv (laplacian.jl:24-71, dst-inversion.jl:39-50)
// as a dense sine-matrix product on the MFMA units, in FP32 or in split BF16
// (three products of high and low parts) with FP32 accumulation.  The
// eigenvalue scaling between the transforms, the residuals that feed them
// and the updates they produce stay in FP64.
//
// Fields between the passes are FP32 in a padded layout: entry (i, j)
// (0-based) of a batch entry at i + P1 j, P1 = n1 rounded up to 64 rows and
// P2 = n2 rounded up to 64 columns, every pad entry zero.  The sine matrices
// are P x P with zero pads, S[k][i] = 2 sin(pi (k+1)(i+1) / (n+1)) (RODFT00's
// unnormalised kernel), symmetric.
#include "mixed.h"

namespace ibpm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ __bf16 to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    unsigned short s = (unsigned short)(u >> 16);
    return __builtin_bit_cast(__bf16, s);
}
__device__ __forceinline__ float bf16_to_f(__bf16 b)
{
    unsigned short s = __builtin_bit_cast(unsigned short, b);
    return __uint_as_float(((uint32_t)s) << 16);
}

// ---------------------------------------------------------------------------
// FP64 residuals and conversions into the padded FP32 layout
// ---------------------------------------------------------------------------
__device__ __forceinline__ double at0(const double* __restrict__ f, int n1, int n2, int i, int j)
{
    return (i >= 0 && i < n1 && j >= 0 && j < n2) ? f[i + (long)n1 * j] : 0.0;
}

// 5-point operator with zero Dirichlet values: 4 u - (neighbours)
__device__ __forceinline__ double lap5(const double* __restrict__ f, int n1, int n2, int i, int j)
{
    const double c = f[i + (long)n1 * j];
    const double w = at0(f, n1, n2, i - 1, j), e = at0(f, n1, n2, i + 1, j);
    const double s = at0(f, n1, n2, i, j - 1), n = at0(f, n1, n2, i, j + 1);
    return 4.0 * c - (((w + e) + s) + n);
}

// apply_bc!(0, get_bc!(coarse), 1) at (i, j), 0-based (mg-utils.jl:7-60)
__device__ __forceinline__ double bc_term(const double* __restrict__ coarse, const GridDev& g, int i, int j)
{
    double t = 0.0;
    if (j == 0) t += get_bc_val(coarse, g, g.BOT + 2 + i);
    if (j == g.n2 - 1) t += get_bc_val(coarse, g, g.TOP + 2 + i);
    if (i == 0) t += get_bc_val(coarse, g, g.LEFT + 2 + j);
    if (i == g.n1 - 1) t += get_bc_val(coarse, g, g.RIGHT + 2 + j);
    return t;
}

__global__ void __launch_bounds__(256) k_mx_resid(MxResidArgs a)
{
    const int b = blockIdx.y;
    const long nP = (long)a.P1 * a.P2;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= nP) return;
    const int i = (int)(e % a.P1), j = (int)(e / a.P1);
    const GridDev& g = a.g;
    const int lev = a.lev0 + b;  // 0-based grid level
    double r = 0.0;
    if (i < g.n1 && j < g.n2) {
        const long p = i + (long)g.n1 * j;
        const double* G = a.G + (long)b * a.fstride;
        if (a.mode == MX_RES_TRIAL) {
            // A Γ - Ainv^{-1} Γ + β1 dt N1 + β2 dt N0 + rhsbc = -2 qc L Γ + ...
            const double hc = level_hc(g.h, lev);
            const double qc = a.dt / (2.0 * a.Re * hc * hc);
            r = ((-2.0 * qc) * lap5(G, g.n1, g.n2, i, j) + a.c1 * a.N1[(long)b * a.fstride + p]) +
                a.c2 * a.N0[(long)b * a.fstride + p];
            if (lev < a.mg - 1) {
                const double fac = 0.25 * a.dt / (a.Re * hc * hc);
                r += fac * bc_term(G + a.fstride, g, i, j);
            }
        } else if (a.mode == MX_RES_TRIAL2) {
            // ((Γ - Γs) - qc (L Γ + L Γs)) + β1 dt N1 + β2 dt N0 + rhsbc
            const double* Gs = a.G2 + (long)b * a.fstride;
            const double hc = level_hc(g.h, lev);
            const double qc = a.dt / (2.0 * a.Re * hc * hc);
            r = (((G[p] - Gs[p]) - qc * (lap5(G, g.n1, g.n2, i, j) + lap5(Gs, g.n1, g.n2, i, j))) +
                 a.c1 * a.N1[(long)b * a.fstride + p]) +
                a.c2 * a.N0[(long)b * a.fstride + p];
            if (lev < a.mg - 1) {
                const double fac = 0.25 * a.dt / (a.Re * hc * hc);
                r += fac * bc_term(G + a.fstride, g, i, j);
            }
        } else if (a.mode == MX_RES_PROJ2) {
            // g - (c + qc L c)
            const double* c = a.G2 + (long)b * a.fstride;
            const double hc = level_hc(g.h, lev);
            const double qc = a.dt / (2.0 * a.Re * hc * hc);
            r = G[p] - (c[p] + qc * lap5(c, g.n1, g.n2, i, j));
        } else if (a.mode == MX_RES_V2F) {
            // Γ + apply_bc!(get_bc!(ψ of the coarser level)) - L ψ
            const double* psi = a.psi + (long)b * a.fstride;
            r = G[p] - lap5(psi, g.n1, g.n2, i, j);
            if (lev < a.mg - 1) r += bc_term(psi + a.fstride, g, i, j);
        } else {  // MX_RES_COPY
            r = G[p];
        }
        r *= a.lscale[lev];
    }
    a.out[(long)b * a.ostride + e] = (float)r;
}

void launch_mx_resid(const MxResidArgs& a, int nbatch, hipStream_t s)
{
    const long nP = (long)a.P1 * a.P2;
    dim3 grid((unsigned)((nP + 255) / 256), (unsigned)nbatch);
    hipLaunchKernelGGL(k_mx_resid, grid, dim3(256), 0, s, a);
}

__global__ void __launch_bounds__(256) k_mx_axpy(double* __restrict__ y, const double* __restrict__ x, double a, long n)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) y[e] += a * x[e];
}

void launch_mx_axpy(double* y, const double* x, double a, long n, hipStream_t s)
{
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mx_axpy, dim3((unsigned)blocks), dim3(256), 0, s, y, x, a, n);
}

// ---------------------------------------------------------------------------
// Sine-matrix GEMM on MFMA.  Out[m, n] = sum_k A[m, k] B[k, n] over one
// 64 x 64 output tile per work-group, four waves of 32 x 32 (2 x 2 blocks of
// 16 x 16).  DIR 0 (x): A = S1 (sine), B = In (k = i contiguous); DIR 1 (y):
// A = In (m = i contiguous), B = S2.  K runs over the padded length in
// chunks of 32 staged through LDS, both operands k-contiguous per row.
//   PREC 0: v_mfma_f32_16x16x4_f32, FP32 operands;
//   PREC 1: v_mfma_f32_16x16x32_bf16 on x = hi + lo splits (hi = bf16(x),
//           lo = bf16(x - hi)): acc += hi*hi + hi*lo + lo*hi.
// Lane layouts (16 x 16 output block, lane l): A row l % 16, B column l % 16,
// k group l / 16; the accumulator holds rows 4 (l / 16) + r, r < 4, of
// column l % 16.
// ---------------------------------------------------------------------------
constexpr int MT = 64, NT = 64, KC = 32;
constexpr int FP = KC + 4;   // FP32 LDS row pitch (floats): 16-byte rows, conflict-free b128 reads
constexpr int BP = KC + 8;   // BF16 LDS row pitch (bf16): 80-byte rows

template <int PREC>
struct MxLds;
template <>
struct MxLds<0> {
    float a[MT][FP];
    float b[NT][FP];
};
template <>
struct MxLds<1> {
    __bf16 ah[MT][BP];
    __bf16 al[MT][BP];
    __bf16 bh[NT][BP];
    __bf16 bl[NT][BP];
};

// stage 8 consecutive-k FP32 values of one row into the tile (row r, k offset c8)
template <int PREC>
__device__ __forceinline__ void put8(MxLds<PREC>& L, bool isA, int r, int c8, const float* v)
{
    if constexpr (PREC == 0) {
        float* d = isA ? &L.a[r][c8] : &L.b[r][c8];
        *reinterpret_cast<float4*>(d) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4*>(d + 4) = make_float4(v[4], v[5], v[6], v[7]);
    } else {
        bf16x8 h, l;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const __bf16 hb = to_bf16(v[t]);
            h[t] = hb;
            l[t] = to_bf16(v[t] - bf16_to_f(hb));
        }
        *reinterpret_cast<bf16x8*>(isA ? &L.ah[r][c8] : &L.bh[r][c8]) = h;
        *reinterpret_cast<bf16x8*>(isA ? &L.al[r][c8] : &L.bl[r][c8]) = l;
    }
}

// The epilogue of one 4-row piece (rows m..m+3 of column n, entry b at level
// lev) of a product: padded FP32 store, with the FP64 eigenvalue scaling, or
// the FP64 update of the unpadded field.
template <int EPI>
__device__ __forceinline__ void mx_epilogue(const MxGemmArgs& a, int b, int lev, int m, int n, float v0, float v1,
                                            float v2, float v3)
{
    const GridDev& g = a.g;
    const float v[4] = {v0, v1, v2, v3};
    if constexpr (EPI == MX_EPI_STORE || EPI == MX_EPI_SCALE) {
        float o[4] = {v[0], v[1], v[2], v[3]};
        if constexpr (EPI == MX_EPI_SCALE) {
            // FP64 eigenvalue scaling: Δinv scale/Λ, Ainv scale/(1 + Λ dt/(2 Re hc^2))
            const double hc = level_hc(g.h, lev);
            const double qc = a.dt / (2.0 * a.Re * hc * hc);
            const double cyv = n < g.n2 ? a.cy[n] : 0.0;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = m + r;
                double f = 0.0;
                if (i < g.n1 && n < g.n2) {
                    const double lam = (4.0 - 2.0 * a.cx[i]) - 2.0 * cyv;
                    const double d = a.op == MX_OP_DINV ? lam : 1.0 + lam * qc;
                    f = a.scale / d;
                }
                o[r] = (float)((double)v[r] * f);
            }
        }
        float* op = a.out + (long)b * a.ostride + m + (long)a.P1 * n;
        *reinterpret_cast<float4*>(op) = make_float4(o[0], o[1], o[2], o[3]);
    } else {  // FP64 update of the unpadded field: dst = base +/- acc / lscale
        if (n < g.n2) {
            const double us = 1.0 / a.lscale[lev];
            double* dst = a.dst + (long)b * a.fstride;
            const double* base = a.base ? a.base + (long)b * a.fstride : nullptr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = m + r;
                if (i < g.n1) {
                    const long p = i + (long)g.n1 * n;
                    const double d = (double)v[r] * us;
                    const double b0 = base ? base[p] : 0.0;
                    dst[p] = EPI == MX_EPI_ADD ? b0 + d : b0 - d;
                }
            }
        }
    }
}

template <int PREC, int DIR, int EPI>
__global__ void __launch_bounds__(256) k_mx_gemm(MxGemmArgs a)
{
    __shared__ MxLds<PREC> L;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int mt = a.P1 / MT, nt = a.P2 / NT;
    const int lin = xcd_block(blockIdx.x, gridDim.x / NXCD);
    if (lin >= mt * nt) return;
    const int S = a.splitk;              // 1: the whole K; else this work-group's share ks of it
    const int b = blockIdx.y / S, ks = blockIdx.y % S;
    const int m0 = (lin % mt) * MT, n0 = (lin / mt) * NT;
    const float* in = a.in + (long)b * a.istride;
    const int K = DIR == 0 ? a.P1 : a.P2;
    const int kb = ks * (K / S), ke = kb + K / S;  // K / S is a multiple of KC

    f32x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};

    // staging assignment: k-contiguous rows (64 rows x 32 k): thread -> row
    // tid / 4, k offset 8 (tid % 4); m-contiguous input (DIR 1, A): thread ->
    // k = tid / 8, m offset 8 (tid % 8)
    const int sr = tid >> 2, sc = (tid & 3) * 8;
    const int tk = tid >> 3, tm = (tid & 7) * 8;

    for (int k0 = kb; k0 < ke; k0 += KC) {
        float va[8], vb[8];
        // ---- global loads of the chunk (all in flight before any LDS store)
        // sine operand rows (k-contiguous): A rows m (DIR 0) or B rows n (DIR 1)
        const int srow = (DIR == 0 ? m0 : n0) + sr;
        __bf16 sh[8], sl[8];
        if constexpr (PREC == 0) {
            const float* sp = a.sf + (long)srow * K + k0 + sc;
            const float4 u0 = *reinterpret_cast<const float4*>(sp);
            const float4 u1 = *reinterpret_cast<const float4*>(sp + 4);
            float* d = DIR == 0 ? va : vb;
            d[0] = u0.x; d[1] = u0.y; d[2] = u0.z; d[3] = u0.w;
            d[4] = u1.x; d[5] = u1.y; d[6] = u1.z; d[7] = u1.w;
        } else {
            const bf16x8 h = *reinterpret_cast<const bf16x8*>(a.sh + (long)srow * K + k0 + sc);
            const bf16x8 l = *reinterpret_cast<const bf16x8*>(a.sl + (long)srow * K + k0 + sc);
#pragma unroll
            for (int t = 0; t < 8; t++) { sh[t] = h[t]; sl[t] = l[t]; }
        }
        // data operand
        if constexpr (DIR == 0) {  // B[k][n] = In[k + P1 n]: row n, k contiguous
            const float* p = in + (long)(n0 + sr) * a.P1 + k0 + sc;
            const float4 u0 = *reinterpret_cast<const float4*>(p);
            const float4 u1 = *reinterpret_cast<const float4*>(p + 4);
            vb[0] = u0.x; vb[1] = u0.y; vb[2] = u0.z; vb[3] = u0.w;
            vb[4] = u1.x; vb[5] = u1.y; vb[6] = u1.z; vb[7] = u1.w;
        } else {  // A[m][k] = In[m + P1 k]: 8 consecutive m at one k
            const float* p = in + (long)(k0 + tk) * a.P1 + m0 + tm;
            const float4 u0 = *reinterpret_cast<const float4*>(p);
            const float4 u1 = *reinterpret_cast<const float4*>(p + 4);
            va[0] = u0.x; va[1] = u0.y; va[2] = u0.z; va[3] = u0.w;
            va[4] = u1.x; va[5] = u1.y; va[6] = u1.z; va[7] = u1.w;
        }
        __syncthreads();  // the previous chunk's reads are done
        // ---- LDS stores
        if constexpr (PREC == 0) {
            if constexpr (DIR == 0) put8<0>(L, true, sr, sc, va);
            else put8<0>(L, false, sr, sc, vb);
        } else {
            bf16x8 h, l;
#pragma unroll
            for (int t = 0; t < 8; t++) { h[t] = sh[t]; l[t] = sl[t]; }
            if constexpr (DIR == 0) {
                *reinterpret_cast<bf16x8*>(&L.ah[sr][sc]) = h;
                *reinterpret_cast<bf16x8*>(&L.al[sr][sc]) = l;
            } else {
                *reinterpret_cast<bf16x8*>(&L.bh[sr][sc]) = h;
                *reinterpret_cast<bf16x8*>(&L.bl[sr][sc]) = l;
            }
        }
        if constexpr (DIR == 0) {
            put8<PREC>(L, false, sr, sc, vb);
        } else {  // transposed: 8 rows m, one k each
#pragma unroll
            for (int t = 0; t < 8; t++) {
                if constexpr (PREC == 0) {
                    L.a[tm + t][tk] = va[t];
                } else {
                    const __bf16 hb = to_bf16(va[t]);
                    L.ah[tm + t][tk] = hb;
                    L.al[tm + t][tk] = to_bf16(va[t] - bf16_to_f(hb));
                }
            }
        }
        __syncthreads();
        // ---- MFMA over the chunk
        const int r16 = lane & 15, g4 = lane >> 4;
        if constexpr (PREC == 0) {
#pragma unroll
            for (int kk = 0; kk < KC; kk += 16) {
                float4 av[2], bv[2];
#pragma unroll
                for (int x = 0; x < 2; x++)
                    av[x] = *reinterpret_cast<const float4*>(&L.a[wm * 32 + x * 16 + r16][kk + 4 * g4]);
#pragma unroll
                for (int y = 0; y < 2; y++)
                    bv[y] = *reinterpret_cast<const float4*>(&L.b[wn * 32 + y * 16 + r16][kk + 4 * g4]);
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++) {
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x].x, bv[y].x, acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x].y, bv[y].y, acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x].z, bv[y].z, acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[x].w, bv[y].w, acc[x][y], 0, 0, 0);
                    }
            }
        } else {
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int x = 0; x < 2; x++) {
                ah[x] = *reinterpret_cast<const bf16x8*>(&L.ah[wm * 32 + x * 16 + r16][8 * g4]);
                al[x] = *reinterpret_cast<const bf16x8*>(&L.al[wm * 32 + x * 16 + r16][8 * g4]);
            }
#pragma unroll
            for (int y = 0; y < 2; y++) {
                bh[y] = *reinterpret_cast<const bf16x8*>(&L.bh[wn * 32 + y * 16 + r16][8 * g4]);
                bl[y] = *reinterpret_cast<const bf16x8*>(&L.bl[wn * 32 + y * 16 + r16][8 * g4]);
            }
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
                }
        }
    }

    // ---- epilogue: lane holds rows m = mb + 4 (lane / 16) + r of column n = nb + lane % 16
    const int lev = a.lev0 + (a.fixlev ? 0 : b);
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            const int m = m0 + wm * 32 + x * 16 + 4 * (lane >> 4);
            const int n = n0 + wn * 32 + y * 16 + (lane & 15);
            const f32x4 v = acc[x][y];
            if (S > 1) {  // partial product: the reduction applies the epilogue
                float* pp = a.part + ((long)ks * (gridDim.y / S) + b) * a.ostride + m + (long)a.P1 * n;
                *reinterpret_cast<float4*>(pp) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                mx_epilogue<EPI>(a, b, lev, m, n, v[0], v[1], v[2], v[3]);
            }
        }
}

// Split-K reduction: the S partial products of every output entry added in
// the order of their K ranges, then the epilogue.  One thread per 4 rows m of
// a column n of entry b.
template <int EPI>
__global__ void __launch_bounds__(256) k_mx_reduce(MxGemmArgs a, int nbatch)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    const long nq = (long)(a.P1 / 4) * a.P2;
    if (e >= nq) return;
    const int m = 4 * (int)(e % (a.P1 / 4)), n = (int)(e / (a.P1 / 4));
    const long off = m + (long)a.P1 * n;
    float4 v = *reinterpret_cast<const float4*>(a.part + (long)b * a.ostride + off);
    for (int s = 1; s < a.splitk; s++) {
        const float4 u = *reinterpret_cast<const float4*>(a.part + ((long)s * nbatch + b) * a.ostride + off);
        v.x += u.x; v.y += u.y; v.z += u.z; v.w += u.w;
    }
    const int lev = a.lev0 + (a.fixlev ? 0 : b);
    mx_epilogue<EPI>(a, b, lev, m, n, v.x, v.y, v.z, v.w);
}

// Launches with fewer than 4 work-groups per CU (one level of a large grid,
// or small grids) split K over S <= MX_SPLIT_MAX work-groups per output tile
// when K / S stays a multiple of the chunk KC.
constexpr int MX_CUS = 256;

template <int PREC, int DIR, int EPI>
static void launch_g(const MxGemmArgs& a0, int nbatch, hipStream_t s)
{
    MxGemmArgs a = a0;
    const int tiles = (a.P1 / MT) * (a.P2 / NT);
    const int K = DIR == 0 ? a.P1 : a.P2;
    int S = 1;
    if (a.part && tiles * nbatch < 4 * MX_CUS) {
        for (int t = MX_SPLIT_MAX; t > 1; t--) {
            if ((K / KC) % t == 0 && (long)t * nbatch * a.ostride <= a.partcap) {
                S = t;
                break;
            }
        }
    }
    a.splitk = S;
    dim3 grid((unsigned)(NXCD * xcd_nper(tiles)), (unsigned)(nbatch * S));
    hipLaunchKernelGGL((k_mx_gemm<PREC, DIR, EPI>), grid, dim3(256), 0, s, a);
    if (S > 1) {
        const long nq = (long)(a.P1 / 4) * a.P2;
        hipLaunchKernelGGL((k_mx_reduce<EPI>), dim3((unsigned)((nq + 255) / 256), (unsigned)nbatch), dim3(256), 0, s,
                           a, nbatch);
    }
}

template <int PREC, int DIR>
static void launch_gd(const MxGemmArgs& a, int epi, int nbatch, hipStream_t s)
{
    switch (epi) {
    case MX_EPI_STORE: launch_g<PREC, DIR, MX_EPI_STORE>(a, nbatch, s); break;
    case MX_EPI_SCALE: launch_g<PREC, DIR, MX_EPI_SCALE>(a, nbatch, s); break;
    case MX_EPI_ADD: launch_g<PREC, DIR, MX_EPI_ADD>(a, nbatch, s); break;
    default: launch_g<PREC, DIR, MX_EPI_SUB>(a, nbatch, s); break;
    }
}

void launch_mx_gemm(int prec, int dir, int epi, const MxGemmArgs& a, int nbatch, hipStream_t s)
{
    if (prec == 0) {
        if (dir == 0) launch_gd<0, 0>(a, epi, nbatch, s);
        else launch_gd<0, 1>(a, epi, nbatch, s);
    } else {
        if (dir == 0) launch_gd<1, 0>(a, epi, nbatch, s);
        else launch_gd<1, 1>(a, epi, nbatch, s);
    }
}

}  // namespace ibpm
