// adjoint.hip -- kernels of the tangent-linear and adjoint advance! steps.
// One thread per output entry throughout: every transpose is written as a
// gather over the forward operator's uses of that entry, so no kernel needs
// atomics and every result is deterministic.  Index comments are Julia
// 1-based, as in the reference files cited.
#include "adjoint.h"

namespace ibpm {

namespace {

struct Idx {
    const GridDev& g;
    __device__ long gx(int i, int j) const { return (i - 1) + (long)(g.nx + 1) * (j - 1); }  // qx[i,j]
    __device__ long gy(int i, int j) const { return g.nu + (i - 1) + (long)g.nx * (j - 1); }  // qy[i,j]
    __device__ long gm(int i, int j) const { return (i - 1) + (long)g.n1 * (j - 1); }        // Γ[i,j]
};

// Γ extended by its boundary ring (direct_product_loops!, nonlinear.jl:73-114):
// Γe[a, b] for a = 0..nx, b = 0..ny; interior a in 1..nx-1, b in 1..ny-1 is
// Γ; the ring is 0.25 get_bc!(Γc): Γe[i-1, 0] = Γbc[B+i], Γe[i-1, ny] =
// Γbc[T+i], Γe[0, j-1] = Γbc[L+j], Γe[nx, j-1] = Γbc[R+j]
__device__ __forceinline__ double gext(const double* __restrict__ G, const double* __restrict__ Gc, const GridDev& g,
                                       int a, int b)
{
    if (a >= 1 && a <= g.n1 && b >= 1 && b <= g.n2) return G[(a - 1) + (long)g.n1 * (b - 1)];
    if (!Gc) return 0.0;
    if (b == 0 && a >= 1 && a <= g.n1) return 0.25 * get_bc_val(Gc, g, g.BOT + a + 1);
    if (b == g.ny && a >= 1 && a <= g.n1) return 0.25 * get_bc_val(Gc, g, g.TOP + a + 1);
    if (a == 0 && b >= 1 && b <= g.n2) return 0.25 * get_bc_val(Gc, g, g.LEFT + b + 1);
    if (a == g.nx && b >= 1 && b <= g.n2) return 0.25 * get_bc_val(Gc, g, g.RIGHT + b + 1);
    return 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// avg_flux and its transpose
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_avg(const double* __restrict__ q, const double* __restrict__ q0,
                                                 double* __restrict__ Q, GridDev g)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= g.nq) return;
    auto v = [&](long k) -> double { return q0 ? q[k] + q0[k] : q[k]; };
    double r = 0.0;
    if (p < g.nu) {  // Qx[i,j] = q[i,j] + q[i,j-1], j = 2..ny
        const int j = (int)(p / (g.nx + 1)) + 1;
        if (j >= 2) r = v(p) + v(p - (g.nx + 1));
    } else {  // Qy[i,j] = q[i,j] + q[i-1,j], i = 2..nx
        const int i = (int)((p - g.nu) % g.nx) + 1;
        if (i >= 2) r = v(p) + v(p - 1);
    }
    Q[p] = r;
}

void launch_adj_avg_flux(const double* q, const double* q0, double* Q, const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_avg, dim3((unsigned)((g.nq + 255) / 256)), dim3(256), 0, s, q, q0, Q, g);
}

__global__ void __launch_bounds__(256) k_adj_avg_T(const double* __restrict__ Qb, double* __restrict__ qb, int acc,
                                                   GridDev g)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= g.nq) return;
    double r = 0.0;
    if (p < g.nu) {  // q[i,j] feeds Qx[i,j] (j >= 2) and Qx[i,j+1] (j+1 <= ny)
        const int j = (int)(p / (g.nx + 1)) + 1;
        if (j >= 2) r += Qb[p];
        if (j + 1 <= g.ny) r += Qb[p + (g.nx + 1)];
    } else {  // q[i,j] feeds Qy[i,j] (i >= 2) and Qy[i+1,j] (i+1 <= nx)
        const int i = (int)((p - g.nu) % g.nx) + 1;
        if (i >= 2) r += Qb[p];
        if (i + 1 <= g.nx) r += Qb[p + 1];
    }
    qb[p] = acc ? qb[p] + r : r;
}

void launch_adj_avg_flux_T(const double* Qb, double* qb, bool acc, const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_avg_T, dim3((unsigned)((g.nq + 255) / 256)), dim3(256), 0, s, Qb, qb, acc ? 1 : 0, g);
}

// ---------------------------------------------------------------------------
// direct product and its transposes
//   fqx[i,j] = Qy[i,j+1] Γe[i-1,j] + Qy[i,j] Γe[i-1,j-1]        i = 2..nx, j = 1..ny
//   fqy[i,j] = -(Qx[i+1,j] Γe[i,j-1] + Qx[i,j] Γe[i-1,j-1])     i = 1..nx, j = 2..ny
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_dprod(const double* __restrict__ Q, const double* __restrict__ G,
                                                   const double* __restrict__ Gc, double* __restrict__ fq, int acc,
                                                   GridDev g)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= g.nq) return;
    const Idx x{g};
    double r = 0.0;
    if (p < g.nu) {
        const int i = (int)(p % (g.nx + 1)) + 1, j = (int)(p / (g.nx + 1)) + 1;
        if (i >= 2 && i <= g.nx)
            r = Q[x.gy(i, j + 1)] * gext(G, Gc, g, i - 1, j) + Q[x.gy(i, j)] * gext(G, Gc, g, i - 1, j - 1);
    } else {
        const long pp = p - g.nu;
        const int i = (int)(pp % g.nx) + 1, j = (int)(pp / g.nx) + 1;
        if (j >= 2 && j <= g.ny)
            r = -(Q[x.gx(i + 1, j)] * gext(G, Gc, g, i, j - 1) + Q[x.gx(i, j)] * gext(G, Gc, g, i - 1, j - 1));
    }
    fq[p] = acc ? fq[p] + r : r;
}

void launch_adj_dprod(const double* Q, const double* G, const double* Gc, double* fq, bool acc, const GridDev& g,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_dprod, dim3((unsigned)((g.nq + 255) / 256)), dim3(256), 0, s, Q, G, Gc, fq,
                       acc ? 1 : 0, g);
}

// Q̄: Qy[i,m] appears in fqx[i,m-1] and fqx[i,m], both times with Γe[i-1,m-1];
// Qx[i,m] appears in fqy[i-1,m] and fqy[i,m], both times with -Γe[i-1,m-1]
__global__ void __launch_bounds__(256) k_adj_dprod_TQ(const double* __restrict__ fb, const double* __restrict__ G,
                                                      const double* __restrict__ Gc, double* __restrict__ Qb,
                                                      GridDev g)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= g.nq) return;
    const Idx x{g};
    double r = 0.0;
    if (p < g.nu) {  // Qx[i,m], i = 1..nx+1, m = 1..ny; used for m = 2..ny
        const int i = (int)(p % (g.nx + 1)) + 1, m = (int)(p / (g.nx + 1)) + 1;
        if (m >= 2) {
            double s = 0.0;
            if (i - 1 >= 1) s += fb[x.gy(i - 1, m)];
            if (i <= g.nx) s += fb[x.gy(i, m)];
            r = -gext(G, Gc, g, i - 1, m - 1) * s;
        }
    } else {  // Qy[i,m], i = 1..nx, m = 1..ny+1; used for i = 2..nx
        const long pp = p - g.nu;
        const int i = (int)(pp % g.nx) + 1, m = (int)(pp / g.nx) + 1;
        if (i >= 2) {
            double s = 0.0;
            if (m - 1 >= 1) s += fb[x.gx(i, m - 1)];
            if (m <= g.ny) s += fb[x.gx(i, m)];
            r = gext(G, Gc, g, i - 1, m - 1) * s;
        }
    }
    Qb[p] = r;
}

void launch_adj_dprod_TQ(const double* fb, const double* G, const double* Gc, double* Qb, const GridDev& g,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_dprod_TQ, dim3((unsigned)((g.nq + 255) / 256)), dim3(256), 0, s, fb, G, Gc, Qb, g);
}

// Γ̄e[a,b]: from fqx (i = a+1 in 2..nx): Qy[a+1,b+1] (f̄qx[a+1,b] + f̄qx[a+1,b+1]),
// rows b and b+1 taken only where they are in 1..ny; from fqy (j = b+1 in 2..ny):
// -Qx[a+1,b+1] (f̄qy[a,b+1] + f̄qy[a+1,b+1]), columns a and a+1 only in 1..nx
__device__ __forceinline__ double gext_bar(const double* __restrict__ fb, const double* __restrict__ Q,
                                           const GridDev& g, int a, int b)
{
    const Idx x{g};
    double r = 0.0;
    if (a + 1 >= 2 && a + 1 <= g.nx) {
        double s = 0.0;
        if (b >= 1 && b <= g.ny) s += fb[x.gx(a + 1, b)];
        if (b + 1 >= 1 && b + 1 <= g.ny) s += fb[x.gx(a + 1, b + 1)];
        r += Q[x.gy(a + 1, b + 1)] * s;
    }
    if (b + 1 >= 2 && b + 1 <= g.ny) {
        double s = 0.0;
        if (a >= 1 && a <= g.nx) s += fb[x.gy(a, b + 1)];
        if (a + 1 >= 1 && a + 1 <= g.nx) s += fb[x.gy(a + 1, b + 1)];
        r -= Q[x.gx(a + 1, b + 1)] * s;
    }
    return r;
}

__global__ void __launch_bounds__(256) k_adj_dprod_TG(const double* __restrict__ fb, const double* __restrict__ Q,
                                                      double* __restrict__ Gb, double* __restrict__ ringb, GridDev g)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < g.nG) {
        const int a = (int)(e % g.n1) + 1, b = (int)(e / g.n1) + 1;
        Gb[e] += gext_bar(fb, Q, g, a, b);
    } else if (e < g.nG + g.nbc) {
        // ring entry k (1-based) of the boundary vector
        const int k = (int)(e - g.nG) + 1;
        double r = 0.0;
        if (k > g.LEFT && k <= g.RIGHT) {  // Γbc[L+j] = Γe[0, j-1], j = 2..ny
            const int j = k - g.LEFT;
            if (j >= 2 && j <= g.ny) r = gext_bar(fb, Q, g, 0, j - 1);
        } else if (k > g.RIGHT && k <= g.BOT) {  // Γbc[R+j] = Γe[nx, j-1]
            const int j = k - g.RIGHT;
            if (j >= 2 && j <= g.ny) r = gext_bar(fb, Q, g, g.nx, j - 1);
        } else if (k > g.BOT && k <= g.TOP) {  // Γbc[B+i] = Γe[i-1, 0], i = 2..nx
            const int i = k - g.BOT;
            if (i >= 2 && i <= g.nx) r = gext_bar(fb, Q, g, i - 1, 0);
        } else if (k > g.TOP) {  // Γbc[T+i] = Γe[i-1, ny]
            const int i = k - g.TOP;
            if (i >= 2 && i <= g.nx) r = gext_bar(fb, Q, g, i - 1, g.ny);
        }
        ringb[k - 1] = r;
    }
}

void launch_adj_dprod_TG(const double* fb, const double* Q, double* Gb, double* ringb, const GridDev& g,
                         hipStream_t s)
{
    const long n = g.nG + g.nbc;
    hipLaunchKernelGGL(k_adj_dprod_TG, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fb, Q, Gb, ringb, g);
}

// ---------------------------------------------------------------------------
// rot! and the single-grid curl!
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_rot(const double* __restrict__ q, double* __restrict__ G, double scale,
                                                 int acc, GridDev g)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.nG) return;
    const Idx x{g};
    const int i = (int)(e % g.n1) + 2, j = (int)(e / g.n1) + 2;  // rot! indices 2..nx, 2..ny
    const double a = q[x.gx(i, j - 1)] - q[x.gx(i, j)];
    const double c = q[x.gy(i, j)] - q[x.gy(i - 1, j)];
    const double r = scale * (a + c);
    G[e] = acc ? G[e] + r : r;
}

void launch_adj_rot(const double* q, double* G, double scale, bool acc, const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_rot, dim3((unsigned)((g.nG + 255) / 256)), dim3(256), 0, s, q, G, scale, acc ? 1 : 0, g);
}

__global__ void __launch_bounds__(256) k_adj_curl_sg(const double* __restrict__ psi, double* __restrict__ q,
                                                     double scale, GridDev g)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= g.nq) return;
    auto PS = [&](int a, int b) -> double {
        return (a >= 1 && a <= g.n1 && b >= 1 && b <= g.n2) ? psi[(a - 1) + (long)g.n1 * (b - 1)] : 0.0;
    };
    double r = 0.0;
    if (p < g.nu) {  // qx[i,j] = ψ[i-1,j] - ψ[i-1,j-1], i = 2..nx
        const int i = (int)(p % (g.nx + 1)) + 1, j = (int)(p / (g.nx + 1)) + 1;
        if (i >= 2 && i <= g.nx) r = PS(i - 1, j) - PS(i - 1, j - 1);
    } else {  // qy[i,j] = ψ[i-1,j-1] - ψ[i,j-1], j = 2..ny
        const long pp = p - g.nu;
        const int i = (int)(pp % g.nx) + 1, j = (int)(pp / g.nx) + 1;
        if (j >= 2 && j <= g.ny) r = PS(i - 1, j - 1) - PS(i, j - 1);
    }
    q[p] = scale * r;
}

void launch_adj_curl_sg(const double* psi, double* q, double scale, const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_curl_sg, dim3((unsigned)((g.nq + 255) / 256)), dim3(256), 0, s, psi, q, scale, g);
}

// ---------------------------------------------------------------------------
// boundary vectors
// ---------------------------------------------------------------------------
// The ψbc part of the multigrid curl! (curl.jl:81-116) per boundary entry k:
//   LEFT  k = L+m: +q̄x(1, m-1) (m >= 2), -q̄x(1, m) (m <= ny), +q̄y(1, m) (2 <= m <= ny)
//   RIGHT k = R+m: +q̄x(nx+1, m-1) (m >= 2), -q̄x(nx+1, m) (m <= ny), -q̄y(nx, m) (2 <= m <= ny)
//   BOT   k = B+m: -q̄x(m+1, 1) (1 <= m <= nx-1), +q̄y(m, 1) (m <= nx), -q̄y(m-1, 1) (m >= 2)
//   TOP   k = T+m: +q̄x(m, ny) (2 <= m <= nx), +q̄y(m, ny+1) (m <= nx), -q̄y(m-1, ny+1) (m >= 2)
__global__ void __launch_bounds__(256) k_adj_curl_bc_T(const double* __restrict__ qb, double* __restrict__ bcb,
                                                       GridDev g)
{
    const int k = blockIdx.x * 256 + threadIdx.x + 1;
    if (k > g.nbc) return;
    const Idx x{g};
    const int nx = g.nx, ny = g.ny;
    double r = 0.0;
    if (k <= g.RIGHT) {
        const int m = k - g.LEFT;
        if (m >= 2) r += qb[x.gx(1, m - 1)];
        if (m <= ny) r -= qb[x.gx(1, m)];
        if (m >= 2 && m <= ny) r += qb[x.gy(1, m)];
    } else if (k <= g.BOT) {
        const int m = k - g.RIGHT;
        if (m >= 2) r += qb[x.gx(nx + 1, m - 1)];
        if (m <= ny) r -= qb[x.gx(nx + 1, m)];
        if (m >= 2 && m <= ny) r -= qb[x.gy(nx, m)];
    } else if (k <= g.TOP) {
        const int m = k - g.BOT;
        if (m >= 1 && m <= nx - 1) r -= qb[x.gx(m + 1, 1)];
        if (m <= nx) r += qb[x.gy(m, 1)];
        if (m >= 2) r -= qb[x.gy(m - 1, 1)];
    } else {
        const int m = k - g.TOP;
        if (m >= 2 && m <= nx) r += qb[x.gx(m, ny)];
        if (m <= nx) r += qb[x.gy(m, ny + 1)];
        if (m >= 2) r -= qb[x.gy(m - 1, ny + 1)];
    }
    bcb[k - 1] = r;
}

void launch_adj_curl_bc_T(const double* qb, double* bcb, const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_curl_bc_T, dim3((unsigned)((g.nbc + 255) / 256)), dim3(256), 0, s, qb, bcb, g);
}

// apply_bc! transposed: bc̄[B+1+i] = fac r̄[i,1], bc̄[T+1+i] = fac r̄[i,ny-1]
// (i = 1..nx-1), bc̄[L+1+j] = fac r̄[1,j], bc̄[R+1+j] = fac r̄[nx-1,j]
// (j = 1..ny-1); every other entry 0
__global__ void __launch_bounds__(256) k_adj_apply_bc_T(const double* __restrict__ rb, double fac,
                                                        double* __restrict__ bcb, GridDev g)
{
    const int k = blockIdx.x * 256 + threadIdx.x + 1;
    if (k > g.nbc) return;
    const Idx x{g};
    double r = 0.0;
    if (k <= g.RIGHT) {
        const int j = k - g.LEFT - 1;
        if (j >= 1 && j <= g.n2) r = rb[x.gm(1, j)];
    } else if (k <= g.BOT) {
        const int j = k - g.RIGHT - 1;
        if (j >= 1 && j <= g.n2) r = rb[x.gm(g.n1, j)];
    } else if (k <= g.TOP) {
        const int i = k - g.BOT - 1;
        if (i >= 1 && i <= g.n1) r = rb[x.gm(i, 1)];
    } else {
        const int i = k - g.TOP - 1;
        if (i >= 1 && i <= g.n1) r = rb[x.gm(i, g.n2)];
    }
    bcb[k - 1] = fac * r;
}

void launch_adj_apply_bc_T(const double* rb, double fac, double* bcb, const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_apply_bc_T, dim3((unsigned)((g.nbc + 255) / 256)), dim3(256), 0, s, rb, fac, bcb, g);
}

// get_bc! transposed onto the coarser level: each coarse point on the rows
// ny/4, 3ny/4 (columns nx/4 .. 3nx/4) or the columns nx/4, 3nx/4 (rows ny/4
// .. 3ny/4) gathers the boundary entries that sample it.  Along a side with
// samples c_m = r[base + m], m = 0..n/2, entry 2m+1 is c_m and entry 2m is
// (c_m + c_{m-1})/2, so c_m receives bc̄[2m+1] + (bc̄[2m] + bc̄[2m+2])/2.
__device__ __forceinline__ double side_T(const double* __restrict__ bcb, int off, int m, int half)
{
    double r = bcb[off + 2 * m + 1 - 1];
    if (m >= 1) r += 0.5 * bcb[off + 2 * m - 1];
    if (m + 1 <= half) r += 0.5 * bcb[off + 2 * m + 2 - 1];
    return r;
}

__global__ void __launch_bounds__(256) k_adj_get_bc_T(const double* __restrict__ bcb, double fac,
                                                      double* __restrict__ rcb, GridDev g)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int hx = g.nx / 2, hy = g.ny / 2;
    const int nrow = hx + 1, ncol = hy - 1;  // columns without the corner rows
    const Idx x{g};
    if (t < 2 * nrow) {  // rows ny/4 (BOT) and 3ny/4 (TOP), m = 0..nx/2
        const int top = t >= nrow, m = t - (top ? nrow : 0);
        const int i = g.nx / 4 + m, j = top ? 3 * g.ny / 4 : g.ny / 4;
        double r = side_T(bcb, top ? g.TOP : g.BOT, m, hx);
        // the row ends are also the columns' end samples (m' = 0 or ny/2)
        if (m == 0 || m == hx) {
            const int off = m == 0 ? g.LEFT : g.RIGHT;
            r += side_T(bcb, off, top ? hy : 0, hy);
        }
        rcb[x.gm(i, j)] += fac * r;
    } else if (t < 2 * nrow + 2 * ncol) {  // columns nx/4 (LEFT) and 3nx/4 (RIGHT), m = 1..ny/2-1
        const int u = t - 2 * nrow;
        const int right = u >= ncol, m = (u - (right ? ncol : 0)) + 1;
        const int i = right ? 3 * g.nx / 4 : g.nx / 4, j = g.ny / 4 + m;
        rcb[x.gm(i, j)] += fac * side_T(bcb, right ? g.RIGHT : g.LEFT, m, hy);
    }
}

void launch_adj_get_bc_T(const double* bcb, double fac, double* rcb, const GridDev& g, hipStream_t s)
{
    const int n = 2 * (g.nx / 2 + 1) + 2 * (g.ny / 2 - 1);
    hipLaunchKernelGGL(k_adj_get_bc_T, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bcb, fac, rcb, g);
}

// ---------------------------------------------------------------------------
// coarsify! transposed.  Fine point (a, b) is covered by the coarse centres
// i = i0 + 2k (ic = ic0 + k, k < ci) with |a - i| <= 1, weight 1 at the centre
// and 1/2 per unit offset in each direction.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_coarsify_T(double* __restrict__ Gfb, const double* __restrict__ Gcb,
                                                        CoarsifyArgs r)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const GridDev& g = r.g;
    if (e >= g.nG) return;
    const int a = (int)(e % g.n1) + 1, b = (int)(e / g.n1) + 1;
    double s = 0.0;
    for (int di = -1; di <= 1; di++) {
        const int ci = a - di;  // the coarse centre's fine index
        if (((ci - r.i0) & 1) != 0) continue;
        const int k = (ci - r.i0) / 2;
        if (ci < r.i0 || k >= r.ci) continue;
        const double wi = di == 0 ? 1.0 : 0.5;
        for (int dj = -1; dj <= 1; dj++) {
            const int cj = b - dj;
            if (((cj - r.j0) & 1) != 0) continue;
            const int l = (cj - r.j0) / 2;
            if (cj < r.j0 || l >= r.cj) continue;
            const double wj = dj == 0 ? 1.0 : 0.5;
            s += (wi * wj) * Gcb[(r.ic0 + k - 1) + (long)g.n1 * (r.jc0 + l - 1)];
        }
    }
    Gfb[e] += s;
}

__global__ void __launch_bounds__(256) k_adj_zero_block(double* __restrict__ Gcb, CoarsifyArgs r)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)r.ci * r.cj) return;
    const int k = (int)(e % r.ci), l = (int)(e / r.ci);
    Gcb[(r.ic0 + k - 1) + (long)r.g.n1 * (r.jc0 + l - 1)] = 0.0;
}

void launch_adj_coarsify_T(double* Gfb, double* Gcb, const CoarsifyArgs& r, hipStream_t s)
{
    if (r.ci <= 0 || r.cj <= 0) return;
    hipLaunchKernelGGL(k_adj_coarsify_T, dim3((unsigned)((r.g.nG + 255) / 256)), dim3(256), 0, s, Gfb,
                       (const double*)Gcb, r);
    const long nb = (long)r.ci * r.cj;
    hipLaunchKernelGGL(k_adj_zero_block, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, Gcb, r);
}

// ---------------------------------------------------------------------------
// small vector kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_axpy(double* __restrict__ y, const double* __restrict__ x, double a,
                                                  long n)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) y[e] += a * x[e];
}

void launch_adj_axpy(double* y, const double* x, double a, long n, hipStream_t s)
{
    long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) return;
    hipLaunchKernelGGL(k_adj_axpy, dim3((unsigned)b), dim3(256), 0, s, y, x, a, n);
}

__global__ void __launch_bounds__(256) k_adj_add_halves(double* __restrict__ x, int n, double a, double b)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < n) x[e] += e < n / 2 ? a : b;
}

void launch_adj_add_halves(double* x, int n, double a, double b, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_adj_add_halves, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, a, b);
}

__global__ void __launch_bounds__(256) k_adj_gemv(const double* __restrict__ M, const double* __restrict__ x,
                                                  double* __restrict__ y, int n)
{
    // one wavefront per row
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* mr = M + (long)row * n;
    double s = 0.0;
    for (int k = lane; k < n; k += 64) s += mr[k] * x[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[row] = s;
}

void launch_adj_gemv(const double* M, const double* x, double* y, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_adj_gemv, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, M, x, y, n);
}

// ---------------------------------------------------------------------------
// reductions (fixed order: per-group partial sums, then one group)
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS_PER_LEVEL = 64;

int adj_q0_reduce_parts(const GridDev& g, int mg)
{
    (void)g;
    return RED_BLOCKS_PER_LEVEL * mg;
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(256) k_adj_q0_part(const double* __restrict__ q0b, double offx, double offy,
                                                     double* __restrict__ part, GridDev g)
{
    __shared__ double sh[4];
    const int lev = blockIdx.y;
    const double hc = g.h * (double)(1L << lev);
    const double* qb = q0b + (long)lev * g.nq;
    double sx = 0.0, sy = 0.0, so = 0.0;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < g.nq; p += (long)gridDim.x * 256) {
        const double v = qb[p];
        if (p < g.nu) {
            const int j = (int)(p / (g.nx + 1)) + 1;
            const double y = ((((double)j - 0.5) - (double)g.ny / 2.0) * hc + ((double)g.ny / 2.0) * g.h) - offy;
            sx += v;
            so -= y * v;
        } else {
            const int i = (int)((p - g.nu) % g.nx) + 1;
            const double x = ((((double)i - 0.5) - (double)g.nx / 2.0) * hc + ((double)g.nx / 2.0) * g.h) - offx;
            sy += v;
            so += x * v;
        }
    }
    sx = block_sum(sx, sh);
    sy = block_sum(sy, sh);
    so = block_sum(so, sh);
    if (threadIdx.x == 0) {
        double* o = part + 3L * (lev * gridDim.x + blockIdx.x);
        o[0] = hc * sx;
        o[1] = hc * sy;
        o[2] = hc * so;
    }
}

__global__ void __launch_bounds__(64) k_adj_part_sum(const double* __restrict__ part, int nparts, double* out)
{
    if (threadIdx.x >= 3) return;
    double s = 0.0;
    for (int k = 0; k < nparts; k++) s += part[3L * k + threadIdx.x];
    out[threadIdx.x] = s;
}

void launch_adj_q0_reduce(const double* q0b, int mg, double offx, double offy, double* part, double* out,
                          const GridDev& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_q0_part, dim3(RED_BLOCKS_PER_LEVEL, (unsigned)mg), dim3(256), 0, s, q0b, offx, offy,
                       part, g);
    hipLaunchKernelGGL(k_adj_part_sum, dim3(1), dim3(64), 0, s, (const double*)part, RED_BLOCKS_PER_LEVEL * mg, out);
}

__global__ void __launch_bounds__(256) k_adj_sum_halves(const double* __restrict__ x, int n, double* out2)
{
    __shared__ double sh[4];
    double a = 0.0, b = 0.0;
    for (int e = threadIdx.x; e < n; e += 256) {
        if (e < n / 2) a += x[e];
        else b += x[e];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        out2[0] = a;
        out2[1] = b;
    }
}

void launch_adj_sum_halves(const double* x, int n, double* out2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_sum_halves, dim3(1), dim3(256), 0, s, x, n, out2);
}

}  // namespace ibpm
