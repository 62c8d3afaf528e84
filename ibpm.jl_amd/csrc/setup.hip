// setup.hip -- device kernels of problem setup: the inverse of the Schur
// matrix B (get_Binv's inv(B), laplacian.jl:118) by Gauss-Jordan elimination
// with partial (row) pivoting on the augmented [B | I], in FP64.
#include "kernels.h"

namespace ibpm {

// pivot of column k: the row p >= k of largest |M[p, k]| (first one on ties),
// and M[p, k]; M is n x 2n column-major (leading dimension n)
__global__ void __launch_bounds__(256) k_gj_pivot(const double* __restrict__ M, int n, int k, int* piv, double* pval,
                                                   int* singular, double* __restrict__ f)
{
    __shared__ double bv[256];
    __shared__ int bi[256];
    double best = -1.0;
    int bidx = n;
    for (int i = k + threadIdx.x; i < n; i += 256) {
        const double v = fabs(M[i + (long)n * k]);
        if (v > best) { best = v; bidx = i; }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = bidx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double a = bv[threadIdx.x], b = bv[threadIdx.x + s];
            const int ia = bi[threadIdx.x], ib = bi[threadIdx.x + s];
            if (b > a || (b == a && ib < ia)) { bv[threadIdx.x] = b; bi[threadIdx.x] = ib; }
        }
        __syncthreads();
    }
    const int p = bi[0];
    if (threadIdx.x == 0) {
        *piv = p;
        *pval = p < n ? M[p + (long)n * k] : 0.0;
        if (!(bv[0] > 0.0)) *singular = 1;
    }
    // multipliers of the elimination: column k as it is after the row swap,
    // row k's set to 0 (read here, before k_gj_swap rewrites the column)
    const double* ck = M + (long)n * k;
    for (int i = threadIdx.x; i < n; i += 256) {
        double v = i == p ? ck[k] : ck[i];
        if (i == k) v = 0.0;
        f[i] = v;
    }
}

// swap rows k and p over all 2n columns and divide the new row k by the pivot
__global__ void __launch_bounds__(256) k_gj_swap(double* __restrict__ M, int n, int k, const int* piv,
                                                 const double* pval)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int p = *piv;
    const double pv = *pval;
    if (c < 2 * n && p < n) {
        double* col = M + (long)n * c;
        const double a = col[k], b = col[p];
        col[p] = a;
        col[k] = b / pv;
    }
}

// M[i, c] -= f[i] M[k, c] for every column c > k of the left half and every
// column of the right half (the left columns <= k are done)
__global__ void __launch_bounds__(256) k_gj_elim(double* __restrict__ M, int n, int k, const double* __restrict__ f)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || c >= 2 * n || (c < n && c <= k)) return;
    const double fi = f[i];
    if (fi != 0.0) M[i + (long)n * c] -= fi * M[k + (long)n * c];
}

void launch_gauss_jordan(double* M, int n, int* piv, double* pval, int* singular, double* f, hipStream_t s)
{
    for (int k = 0; k < n; k++) {
        hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(256), 0, s, (const double*)M, n, k, piv, pval, singular, f);
        hipLaunchKernelGGL(k_gj_swap, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0, s, M, n, k,
                           (const int*)piv, (const double*)pval);
        hipLaunchKernelGGL(k_gj_elim, dim3((unsigned)((n + 63) / 64), (unsigned)((2 * n + 3) / 4)), dim3(256), 0, s,
                           M, n, k, (const double*)f);
    }
}

}  // namespace ibpm
