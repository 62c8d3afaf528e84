/*
 * ibpm_oracle.c -- CPU restatement of IBPM.jl's `advance!(state, prob, t)` path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it to produce the
 * numbers the HIP path is compared against.  Nothing in the product library
 * links, loads or calls it.
 *
 * Every function restates one reference function, operation by operation, in
 * the reference's own Julia column-major layouts and (where it fixes rounding)
 * in the reference's own order of floating-point operations.  Index arithmetic
 * is written with 1-based accessor macros so that it reads like the Julia code
 * it follows.  Citations are file:line into the reference tree.
 *
 * The two third-party numerics the reference calls are restated here from
 * their published definitions:
 *   - FFTW's RODFT00 (DST-I, unnormalised) used by `dst_inv!`
 *     (src/fluid-operators/dst-inversion.jl:10-20,39-50): computed through a
 *     complex mixed-radix FFT of the odd extension of length 2(n+1), two real
 *     lines packed per complex transform.
 *   - LAPACK getrf/getri behind `inv(B)` (src/fluid-operators/laplacian.jl:118):
 *     LU with partial pivoting, then column-by-column solves.
 *   - IterativeSolvers.jl 0.9.1's `cg!` behind the MotionFunction force solve
 *     (laplacian.jl:137-138; version from the reference's Manifest.toml:266-270):
 *     its recurrences, initial guess and stopping rule (o_cg).
 *
 * How each part is pinned (tests/test_oracle.py, test_oracle_motion.py,
 * test_reference_data.py): the DST-I against a brute-force sine sum and
 * scipy; the Laplacian inverses against the assembled 5-point operators;
 * adjointness of curl/rot and E/E'; δh moments; closed forms of coarsify,
 * get_bc, apply_bc, the nonlinear term and the base flux; the linearized
 * product against the bilinear cross-term identity; move_body! against rigid
 * motion; the CG against a dense solve; the MotionFunction step against the
 * Static path and, on its first step, against a static body in a freestream
 * and against the MovingGrid frame; end to end against the reference's
 * steady flat-plate CD (10001 steps, 8 digits).  No function is left
 * unpinned.
 *
 * Build: oracle/Makefile (gcc -O2 -fPIC -shared -fopenmp -ffp-contract=off).
 * -ffp-contract=off keeps every a*b+c as two rounded operations, as Julia does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* complex mixed-radix FFT (forward, e^{-2 pi i jk/n})                        */
/* ------------------------------------------------------------------------- */
typedef struct { double re, im; } ocx;

typedef struct {
    int n;
    int nfac;
    int fac[64];
    ocx* tw; /* tw[k] = exp(-2 pi i k / n) */
} offt_plan;

static void offt_plan_init(offt_plan* p, int n)
{
    p->n = n;
    p->nfac = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
    }
    p->tw = (ocx*)malloc(sizeof(ocx) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        p->tw[k].re = (double)cosl(a);
        p->tw[k].im = (double)sinl(a);
    }
}

static void offt_plan_free(offt_plan* p) { free(p->tw); p->tw = NULL; }

/* out[k] = sum_j in[j*is] W_n^{jk};  W_n^{e} = tw[e*tws mod N] */
static void offt_rec(ocx* out, const ocx* in, int n, int is, const offt_plan* p, int d, int tws)
{
    if (n == 1) { out[0] = in[0]; return; }
    /* q*k*tws < r*m*(N/n) = N, so first-stage twiddle indices need no wrap */
    const int r = p->fac[d], m = n / r, N = p->n;
    for (int q = 0; q < r; q++) offt_rec(out + q * m, in + q * is, m, is * r, p, d + 1, tws * r);
    if (r == 2) {
        for (int k = 0; k < m; k++) {
            ocx a = out[k], b = out[m + k], w = p->tw[(long)k * tws];
            ocx t = { b.re * w.re - b.im * w.im, b.re * w.im + b.im * w.re };
            out[k].re = a.re + t.re; out[k].im = a.im + t.im;
            out[m + k].re = a.re - t.re; out[m + k].im = a.im - t.im;
        }
    } else if (r == 4) {
        for (int k = 0; k < m; k++) {
            ocx v[4];
            for (int q = 0; q < 4; q++) {
                ocx a = out[q * m + k], w = p->tw[(long)q * k * tws];
                v[q].re = a.re * w.re - a.im * w.im;
                v[q].im = a.re * w.im + a.im * w.re;
            }
            /* DFT-4 with W_4 = -i */
            ocx s02 = { v[0].re + v[2].re, v[0].im + v[2].im };
            ocx d02 = { v[0].re - v[2].re, v[0].im - v[2].im };
            ocx s13 = { v[1].re + v[3].re, v[1].im + v[3].im };
            ocx d13 = { v[1].re - v[3].re, v[1].im - v[3].im };
            out[k].re = s02.re + s13.re;         out[k].im = s02.im + s13.im;
            out[2 * m + k].re = s02.re - s13.re; out[2 * m + k].im = s02.im - s13.im;
            /* -i * d13 = (d13.im, -d13.re) */
            out[m + k].re = d02.re + d13.im;     out[m + k].im = d02.im - d13.re;
            out[3 * m + k].re = d02.re - d13.im; out[3 * m + k].im = d02.im + d13.re;
        }
    } else {
        ocx v[64], o[64];
        for (int k = 0; k < m; k++) {
            for (int q = 0; q < r; q++) {
                ocx a = out[q * m + k], w = p->tw[(long)q * k * tws];
                v[q].re = a.re * w.re - a.im * w.im;
                v[q].im = a.re * w.im + a.im * w.re;
            }
            for (int s = 0; s < r; s++) {
                double re = 0.0, im = 0.0;
                for (int q = 0; q < r; q++) {
                    ocx w = p->tw[(long)q * s * m * tws % N];
                    re += v[q].re * w.re - v[q].im * w.im;
                    im += v[q].re * w.im + v[q].im * w.re;
                }
                o[s].re = re; o[s].im = im;
            }
            for (int s = 0; s < r; s++) out[k + s * m] = o[s];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* DST-I (FFTW RODFT00): Y_k = 2 sum_{j=0}^{n-1} X_j sin(pi (j+1)(k+1)/(n+1))  */
/* Two real lines a, b per complex FFT of the odd extension z of length      */
/* N = 2(n+1): Z_k = -i Ya_{k-1} + Yb_{k-1}.                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n;           /* line length */
    offt_plan plan;  /* N = 2(n+1) */
} odst_plan;

static void odst_plan_init(odst_plan* p, int n) { p->n = n; offt_plan_init(&p->plan, 2 * (n + 1)); }
static void odst_plan_free(odst_plan* p) { offt_plan_free(&p->plan); }

/* a, b: input lines (stride sa/sb); b may be NULL.  ya, yb: outputs. */
static void odst_pair(const odst_plan* p, const double* a, long sa, const double* b, long sb,
                      double* ya, long sya, double* yb, long syb, ocx* zin, ocx* zout)
{
    const int n = p->n, N = 2 * (n + 1);
    zin[0].re = 0.0; zin[0].im = 0.0;
    zin[n + 1].re = 0.0; zin[n + 1].im = 0.0;
    for (int j = 0; j < n; j++) {
        double ar = a[j * sa], bi = b ? b[j * sb] : 0.0;
        zin[j + 1].re = ar; zin[j + 1].im = bi;
        zin[N - 1 - j].re = -ar; zin[N - 1 - j].im = -bi;
    }
    offt_rec(zout, zin, N, 1, &p->plan, 0, 1);
    for (int k = 0; k < n; k++) {
        ya[k * sya] = -zout[k + 1].im;
        if (yb) yb[k * syb] = zout[k + 1].re;
    }
}

static int o_nthreads = 1;

void oracle_set_threads(int n) { o_nthreads = n > 0 ? n : 1; }

/* Variants of the 2D DST-I, for measuring how far FP64 codes that compute the
 * transform differently drift apart (tests/golden/dst_order_spread.py).
 * Variant 1: the same RODFT00 with one real line per complex FFT (instead of
 * two) and dimension 2 transformed before dimension 1.  Variant 2: the
 * half-length algorithm (below).  Variant 0 (the default) is the oracle's own
 * order. */
static int o_dst_variant = 0;

void oracle_set_dst_variant(int v) { o_dst_variant = v; }

static void o_dst2_variant1(double* y, const double* x, int n1, int n2, const odst_plan* p1, const odst_plan* p2)
{
    const int N1 = 2 * (n1 + 1), N2 = 2 * (n2 + 1);
    const int Nmax = N1 > N2 ? N1 : N2;
    if (y != x) memcpy(y, x, sizeof(double) * (size_t)n1 * (size_t)n2);
    /* dimension 2 first: strided rows, one per transform */
#pragma omp parallel num_threads(o_nthreads)
    {
        ocx* zin = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        ocx* zout = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        double* ta = (double*)malloc(sizeof(double) * (size_t)n2);
        double* ua = (double*)malloc(sizeof(double) * (size_t)n2);
#pragma omp for schedule(static)
        for (int i = 0; i < n1; i++) {
            for (int j = 0; j < n2; j++) ta[j] = y[i + (long)j * n1];
            odst_pair(p2, ta, 1, NULL, 1, ua, 1, NULL, 1, zin, zout);
            for (int j = 0; j < n2; j++) y[i + (long)j * n1] = ua[j];
        }
        free(zin); free(zout); free(ta); free(ua);
    }
    /* then dimension 1: contiguous columns, one per transform */
#pragma omp parallel num_threads(o_nthreads)
    {
        ocx* zin = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        ocx* zout = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        double* ta = (double*)malloc(sizeof(double) * (size_t)n1);
#pragma omp for schedule(static)
        for (int j = 0; j < n2; j++) {
            memcpy(ta, y + (long)j * n1, sizeof(double) * (size_t)n1);
            odst_pair(p1, ta, 1, NULL, 1, y + (long)j * n1, 1, NULL, 1, zin, zout);
        }
        free(zin); free(zout); free(ta);
    }
}

/* 2D RODFT00 over dims [1,2] of the n1 x n2 column-major array x -> y (may alias).
 * Same transform as FFTW.plan_r2r(b, RODFT00, [1,2]) (dst-inversion.jl:10-20). */
/* Variant 2: the half-length DST-I algorithm (even M = n + 1 only): pre-
 * processing y_j = sin(pi j/M)(x_j + x_{M-j}) + (x_j - x_{M-j})/2, a complex
 * FFT of P = M/2 points of z_m = y_{2m} + i y_{2m+1}, real-DFT unpacking into
 * F_k = R_k - i I_k, then Y(2k) = 2 I_k and the odd outputs by the recurrence
 * Y(2k+1) = Y(2k-1) + 2 R_k, Y(1) = R_0.  This is the arithmetic structure of
 * the HIP library's transform (restated here in plain loops); its recurrence
 * sums the FFT's rounding errors along k.  Variant 3: the same with the
 * recurrence summed with Neumaier's compensation (the running sum's rounding
 * error carried in a second term). */
static void o_dst_halflen_line(int n, const double* x, long sx, double* out, long so)
{
    const int comp = o_dst_variant == 3;
    const int M = n + 1, P = M / 2;
    static __thread offt_plan pl;
    static __thread int plP = 0;
    static __thread ocx *zin = NULL, *zout = NULL, *tw = NULL;
    static __thread double* yy = NULL;
    if (plP != P) {
        if (plP) { offt_plan_free(&pl); free(zin); free(zout); free(tw); free(yy); }
        offt_plan_init(&pl, P);
        zin = (ocx*)malloc(sizeof(ocx) * (size_t)P);
        zout = (ocx*)malloc(sizeof(ocx) * (size_t)P);
        tw = (ocx*)malloc(sizeof(ocx) * (size_t)(P + 1));
        yy = (double*)malloc(sizeof(double) * (size_t)M);
        for (int k = 0; k <= P; k++) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)M;
            tw[k].re = (double)cosl(a);
            tw[k].im = (double)sinl(a);
        }
        plP = P;
    }
    for (int j = 0; j < M; j++) {
        double xj = (j >= 1 && j <= n) ? x[(long)(j - 1) * sx] : 0.0;
        double xm = (M - j >= 1 && M - j <= n) ? x[(long)(M - j - 1) * sx] : 0.0;
        double sj = (double)sinl(3.14159265358979323846264338327950288L * (long double)j / (long double)M);
        yy[j] = sj * (xj + xm) + 0.5 * (xj - xm);
    }
    for (int m = 0; m < P; m++) { zin[m].re = yy[2 * m]; zin[m].im = yy[2 * m + 1]; }
    offt_rec(zout, zin, P, 1, &pl, 0, 1);
    double s = 0.0, c = 0.0;
    for (int k = 0; k < P; k++) {
        ocx a = zout[k], b = zout[(P - k) % P];
        /* E = (a + conj b)/2, O = -i (a - conj b)/2, F = E + W^k O */
        double er = 0.5 * (a.re + b.re), ei = 0.5 * (a.im - b.im);
        double dr = 0.5 * (a.re - b.re), di = 0.5 * (a.im + b.im);
        double orr = di, oi = -dr;
        double fr = er + (tw[k].re * orr - tw[k].im * oi);
        double fi = ei + (tw[k].re * oi + tw[k].im * orr);
        double Rk = fr, Ik = -fi;
        if (k == 0) s = Rk;
        else {
            out[(long)(2 * k - 1) * so] = 2.0 * Ik;
            const double y2 = 2.0 * Rk, t = s + y2;
            if (comp) c += fabs(s) >= fabs(y2) ? (s - t) + y2 : (y2 - t) + s;
            s = t;
        }
        out[(long)(2 * k) * so] = s + c;
    }
}

static void o_dst2_variant2(double* y, const double* x, int n1, int n2)
{
    double* t = (double*)malloc(sizeof(double) * (size_t)n1 * (size_t)n2);
#pragma omp parallel for num_threads(o_nthreads) schedule(static)
    for (int j = 0; j < n2; j++) o_dst_halflen_line(n1, x + (long)j * n1, 1, t + (long)j * n1, 1);
#pragma omp parallel for num_threads(o_nthreads) schedule(static)
    for (int i = 0; i < n1; i++) o_dst_halflen_line(n2, t + i, n1, y + i, n1);
    free(t);
}

static void o_dst2(double* y, const double* x, int n1, int n2, const odst_plan* p1, const odst_plan* p2)
{
    if (o_dst_variant == 1) {
        o_dst2_variant1(y, x, n1, n2, p1, p2);
        return;
    }
    if ((o_dst_variant == 2 || o_dst_variant == 3) && (n1 + 1) % 2 == 0 && (n2 + 1) % 2 == 0) {
        o_dst2_variant2(y, x, n1, n2);
        return;
    }
    const int N1 = 2 * (n1 + 1), N2 = 2 * (n2 + 1);
    const int Nmax = N1 > N2 ? N1 : N2;
    /* dimension 1: contiguous columns, paired */
#pragma omp parallel num_threads(o_nthreads)
    {
        ocx* zin = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        ocx* zout = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        double* ta = (double*)malloc(sizeof(double) * (size_t)(n2 > n1 ? n2 : n1));
        double* tb = (double*)malloc(sizeof(double) * (size_t)(n2 > n1 ? n2 : n1));
#pragma omp for schedule(static)
        for (int jp = 0; jp < (n2 + 1) / 2; jp++) {
            int j = 2 * jp;
            const double* a = x + (long)j * n1;
            const double* b = (j + 1 < n2) ? x + (long)(j + 1) * n1 : NULL;
            /* copy inputs first so that y may alias x */
            memcpy(ta, a, sizeof(double) * (size_t)n1);
            if (b) memcpy(tb, b, sizeof(double) * (size_t)n1);
            odst_pair(p1, ta, 1, b ? tb : NULL, 1, y + (long)j * n1, 1,
                      b ? y + (long)(j + 1) * n1 : NULL, 1, zin, zout);
        }
        free(zin); free(zout); free(ta); free(tb);
    }
    /* dimension 2: strided rows, paired */
#pragma omp parallel num_threads(o_nthreads)
    {
        ocx* zin = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        ocx* zout = (ocx*)malloc(sizeof(ocx) * (size_t)Nmax);
        double* ta = (double*)malloc(sizeof(double) * (size_t)n2);
        double* tb = (double*)malloc(sizeof(double) * (size_t)n2);
        double* ua = (double*)malloc(sizeof(double) * (size_t)n2);
        double* ub = (double*)malloc(sizeof(double) * (size_t)n2);
#pragma omp for schedule(static)
        for (int ip = 0; ip < (n1 + 1) / 2; ip++) {
            int i = 2 * ip;
            int hb = (i + 1 < n1);
            for (int j = 0; j < n2; j++) {
                ta[j] = y[i + (long)j * n1];
                if (hb) tb[j] = y[i + 1 + (long)j * n1];
            }
            odst_pair(p2, ta, 1, hb ? tb : NULL, 1, ua, 1, hb ? ub : NULL, 1, zin, zout);
            for (int j = 0; j < n2; j++) {
                y[i + (long)j * n1] = ua[j];
                if (hb) y[i + 1 + (long)j * n1] = ub[j];
            }
        }
        free(zin); free(zout); free(ta); free(tb); free(ua); free(ub);
    }
}

/* ------------------------------------------------------------------------- */
/* grid (src/fluid-domain/fluid-domain.jl:10-25,44-73)                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nG, nu, nv, nq, mg, nbc;
    double offx, offy, len, h;
    int LEFT, RIGHT, BOT, TOP;
} og_grid;

/* make_grid(h, boundary; mg) -- fluid-domain.jl:44-73.  Julia's round() is
 * round-half-to-even, which is nearbyint() under the default rounding mode. */
static void og_make_grid(og_grid* g, double h, double x0, double x1, double y0, double y1, int mg)
{
    g->offx = -x0;
    g->offy = -y0;
    g->len = x1 - x0;
    double ylen = y1 - y0;
    g->nx = (int)nearbyint(g->len / h);
    g->ny = (int)nearbyint(ylen * (double)g->nx / g->len);
    g->nG = (g->nx - 1) * (g->ny - 1);
    g->nu = g->ny * (g->nx + 1);
    g->nv = g->nx * (g->ny + 1);
    g->nq = g->nu + g->nv;
    g->h = g->len / (double)g->nx;
    g->mg = mg;
    g->LEFT = 0;
    g->RIGHT = g->ny + 1;
    g->BOT = 2 * (g->ny + 1);
    g->TOP = 2 * (g->ny + 1) + g->nx + 1;
    g->nbc = 2 * (g->nx + 1) + 2 * (g->ny + 1);
}

/* 1-based accessors in the reference's reshaped views */
#define GM(a, i, j) (a)[((i) - 1) + (long)(nx - 1) * ((j) - 1)]   /* (nx-1) x (ny-1) */
#define QX(a, i, j) (a)[((i) - 1) + (long)(nx + 1) * ((j) - 1)]   /* (nx+1) x ny      */
#define QY(a, i, j) (a)[nu + ((i) - 1) + (long)nx * ((j) - 1)]    /* nx x (ny+1)      */
#define BC(a, k) (a)[(k) - 1]

/* ------------------------------------------------------------------------- */
/* fluid operators                                                            */
/* ------------------------------------------------------------------------- */

/* lap_eigs -- laplacian.jl:14-19.  Lam = -2*(cos(pi*i/nx) + cos(pi*j/ny) - 2) */
static void o_lap_eigs(const og_grid* g, double* Lam)
{
    const int nx = g->nx, ny = g->ny;
    double* cxv = (double*)malloc(sizeof(double) * (size_t)nx);
    double* cyv = (double*)malloc(sizeof(double) * (size_t)ny);
    const long double PI = 3.14159265358979323846264338327950288L;
    /* Julia's pi*(1:n)/n is a twice-precision range: each argument is the
     * double nearest to fl(pi)*i/n; long double reproduces that rounding. */
    const long double fpi = (long double)(double)PI;
    for (int i = 1; i <= nx - 1; i++) cxv[i] = cos((double)(fpi * (long double)i / (long double)nx));
    for (int j = 1; j <= ny - 1; j++) cyv[j] = cos((double)(fpi * (long double)j / (long double)ny));
    for (int j = 1; j <= ny - 1; j++)
        for (int i = 1; i <= nx - 1; i++)
            GM(Lam, i, j) = -2.0 * ((cxv[i] + cyv[j]) - 2.0);
    free(cxv); free(cyv);
}

/* rot! (C') -- curl.jl:22-34 */
static void o_rot(const og_grid* g, double* G, const double* q)
{
    const int nx = g->nx, ny = g->ny, nu = g->nu;
    for (int j = 2; j <= ny; j++)
        for (int i = 2; i <= nx; i++) {
            double a = QX(q, i, j - 1) - QX(q, i, j);
            double b = QY(q, i, j) - QY(q, i - 1, j);
            GM(G, i - 1, j - 1) = a + b;
        }
}

/* curl! single grid (C) -- curl.jl:51-71.  Used only by adjointness checks. */
static void o_curl_sg(const og_grid* g, double* q, const double* psi)
{
    const int nx = g->nx, ny = g->ny, nu = g->nu;
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx; i++) QX(q, i, j) = GM(psi, i - 1, j) - GM(psi, i - 1, j - 1);
    for (int i = 2; i <= nx; i++) QX(q, i, 1) = GM(psi, i - 1, 1);
    for (int i = 2; i <= nx; i++) QX(q, i, ny) = -GM(psi, i - 1, ny - 1);
    for (int j = 2; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) QY(q, i, j) = GM(psi, i - 1, j - 1) - GM(psi, i, j - 1);
    for (int j = 2; j <= ny; j++) QY(q, 1, j) = -GM(psi, 1, j - 1);
    for (int j = 2; j <= ny; j++) QY(q, nx, j) = GM(psi, nx - 1, j - 1);
}

/* curl! multigrid -- curl.jl:81-116 (note the bottom edge reads psibc[(B-1)+i]) */
static void o_curl_mg(const og_grid* g, double* q, const double* psi, const double* psibc)
{
    const int nx = g->nx, ny = g->ny, nu = g->nu;
    const int T = g->TOP, B = g->BOT, L = g->LEFT, R = g->RIGHT;
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx; i++) QX(q, i, j) = GM(psi, i - 1, j) - GM(psi, i - 1, j - 1);
    for (int i = 2; i <= nx; i++) QX(q, i, 1) = GM(psi, i - 1, 1) - BC(psibc, (B - 1) + i);
    for (int i = 2; i <= nx; i++) QX(q, i, ny) = BC(psibc, i + T) - GM(psi, i - 1, ny - 1);
    for (int j = 1; j <= ny; j++) QX(q, 1, j) = BC(psibc, (L + 1) + j) - BC(psibc, L + j);
    for (int j = 1; j <= ny; j++) QX(q, nx + 1, j) = BC(psibc, (R + 1) + j) - BC(psibc, R + j);
    for (int j = 2; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++) QY(q, i, j) = GM(psi, i - 1, j - 1) - GM(psi, i, j - 1);
    for (int j = 2; j <= ny; j++) QY(q, 1, j) = BC(psibc, L + j) - GM(psi, 1, j - 1);
    for (int j = 2; j <= ny; j++) QY(q, nx, j) = GM(psi, nx - 1, j - 1) - BC(psibc, R + j);
    for (int i = 1; i <= nx; i++) QY(q, i, 1) = BC(psibc, B + i) - BC(psibc, (B + 1) + i);
    for (int i = 1; i <= nx; i++) QY(q, i, ny + 1) = BC(psibc, T + i) - BC(psibc, (T + 1) + i);
}

/* get_bc! -- mg-utils.jl:7-32 */
static void o_get_bc(const og_grid* g, double* rbc, const double* r)
{
    const int nx = g->nx, ny = g->ny;
    const int T = g->TOP, B = g->BOT, L = g->LEFT, R = g->RIGHT;
    const int nx4 = nx / 4, ny4 = ny / 4, nx34 = 3 * nx / 4, ny34 = 3 * ny / 4;
    for (int m = 0; m <= nx / 2; m++) {
        int i = nx4 + m, ibc = 1 + 2 * m;
        BC(rbc, B + ibc) = GM(r, i, ny4);
        BC(rbc, T + ibc) = GM(r, i, ny34);
    }
    for (int m = 1; m <= nx / 2; m++) {
        int i = nx4 + m, ibc = 2 * m;
        BC(rbc, B + ibc) = 0.5 * (GM(r, i, ny4) + GM(r, i - 1, ny4));
        BC(rbc, T + ibc) = 0.5 * (GM(r, i, ny34) + GM(r, i - 1, ny34));
    }
    for (int m = 0; m <= ny / 2; m++) {
        int j = ny4 + m, jbc = 1 + 2 * m;
        BC(rbc, L + jbc) = GM(r, nx4, j);
        BC(rbc, R + jbc) = GM(r, nx34, j);
    }
    for (int m = 1; m <= ny / 2; m++) {
        int j = ny4 + m, jbc = 2 * m;
        BC(rbc, L + jbc) = 0.5 * (GM(r, nx4, j) + GM(r, nx4, j - 1));
        BC(rbc, R + jbc) = 0.5 * (GM(r, nx34, j) + GM(r, nx34, j - 1));
    }
}

/* apply_bc! -- mg-utils.jl:43-60 */
static void o_apply_bc(const og_grid* g, double* r, const double* rbc, double fac)
{
    const int nx = g->nx, ny = g->ny;
    const int T = g->TOP, B = g->BOT, L = g->LEFT, R = g->RIGHT;
    for (int i = 1; i <= nx - 1; i++) GM(r, i, 1) += fac * BC(rbc, (B + 1) + i);
    for (int i = 1; i <= nx - 1; i++) GM(r, i, ny - 1) += fac * BC(rbc, (T + 1) + i);
    for (int j = 1; j <= ny - 1; j++) GM(r, 1, j) += fac * BC(rbc, (L + 1) + j);
    for (int j = 1; j <= ny - 1; j++) GM(r, nx - 1, j) += fac * BC(rbc, (R + 1) + j);
}

/* Julia integer division ÷ truncates toward zero, like C's '/'. */
/* coarsify! -- mg-utils.jl:71-106 (zip stops at the shorter range) */
static void o_coarsify(const og_grid* g, double* Gc, const double* G)
{
    const int nx = g->nx, ny = g->ny;
    int j0 = ny / 2 + ((-ny) / 2 + 2), j1 = ny / 2 + (ny / 2 - 2);
    int jc0 = ny / 2 + ((-ny) / 4 + 1), jc1 = ny / 2 + (ny / 4 - 1);
    int i0 = nx / 2 + ((-nx) / 2 + 2), i1 = nx / 2 + (nx / 2 - 2);
    int ic0 = nx / 2 + ((-nx) / 4 + 1), ic1 = nx / 2 + (nx / 4 - 1);
    int nj = j1 >= j0 ? (j1 - j0) / 2 + 1 : 0, njc = jc1 >= jc0 ? jc1 - jc0 + 1 : 0;
    int ni = i1 >= i0 ? (i1 - i0) / 2 + 1 : 0, nic = ic1 >= ic0 ? ic1 - ic0 + 1 : 0;
    int cj = nj < njc ? nj : njc, ci = ni < nic ? ni : nic;
    for (int b = 0; b < cj; b++) {
        int j = j0 + 2 * b, jc = jc0 + b;
        for (int a = 0; a < ci; a++) {
            int i = i0 + 2 * a, ic = ic0 + a;
            double s1 = ((GM(G, i + 1, j) + GM(G, i, j + 1)) + GM(G, i - 1, j)) + GM(G, i, j - 1);
            double s2 = ((GM(G, i + 1, j + 1) + GM(G, i + 1, j - 1)) + GM(G, i - 1, j - 1)) + GM(G, i - 1, j + 1);
            GM(Gc, ic, jc) = (GM(G, i, j) + 0.5 * s1) + 0.25 * s2;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* problem                                                                    */
/* ------------------------------------------------------------------------- */
#define SUPP 6
#define WIN 13 /* 2*SUPP+1 */

typedef struct {
    og_grid g;
    double Re, dt;
    int nbeta;
    double beta[4];
    int nbodies, nbt;
    int* nb;
    double* xb; /* nbt x 2 col-major */
    double* ds;
    int motion; /* 0 = Static, 1 = MovingGrid, 2 = MotionFunction */
    /* MotionFunction: reference points x0 (RigidBody.x0, body-types.jl:13),
     * point velocities ub (nbt x 2 col-major) and the last CG solve */
    double *x0, *ub;
    int cg_iters;
    double cg_res, cg_tol;
    /* precomputed operators */
    double* Lam;     /* mats.Λ */
    double* divA;    /* mg x nG: 1 ./ (1 .- Λ dt/(2 Re hc^2)) */
    double* divAinv; /* mg x nG: 1 .+ Λ dt/(2 Re hc^2) */
    double lap_scale;
    int* body_idx;   /* nbt x 2 */
    double* weight;  /* [k][c][l + 13 m] */
    double* Binv;    /* nf x nf col-major */
    double* XX;      /* nx(ny+1) x mg */
    double* YY;      /* ny(nx+1) x mg */
    odst_plan p1, p2;
    /* WorkingMemory (models.jl:79-106) */
    double *q1, *G1, *q2, *G2, *q3, *q4, *q5, *q6, *G3, *Gbc, *rhsbc, *Gwork, *wdst;
    double *bt_G, *bt_psi, *bt_q; /* get_Binv scratch */
    /* LinearizedIBProblem (problem-types.jl:47-87): base flux QB (avg_flux of
     * the base state, nq x mg) and base circulation ΓB (nG x mg); lin = 1 */
    int lin;
    double *QB, *GB;
} oprob;

/* dst_inv!(x, b, Λ, plan; scale) -- dst-inversion.jl:39-50, wrapped by
 * get_lap_inv -- laplacian.jl:24-41 (copy through b_temp/x_temp). */
/* Optional replacement of dst_inv! for error studies of reduced-precision
 * transforms (the hook receives x, b, Λ, scale, nx-1, ny-1); NULL restores
 * the oracle's own FP64 transform. */
typedef void (*o_lapinv_hook_t)(double* x, const double* b, const double* Lam, double scale, int n1, int n2);
static o_lapinv_hook_t o_lapinv_hook = NULL;
void oracle_set_lapinv_hook(o_lapinv_hook_t f) { o_lapinv_hook = f; }

static void o_dst_inv(const oprob* P, double* x, const double* b, const double* Lam, double scale)
{
    const og_grid* g = &P->g;
    if (o_lapinv_hook) {
        o_lapinv_hook(x, b, Lam, scale, g->nx - 1, g->ny - 1);
        return;
    }
    double* w = P->wdst;
    o_dst2(w, b, g->nx - 1, g->ny - 1, &P->p1, &P->p2);
    for (long k = 0; k < g->nG; k++) w[k] /= Lam[k];
    for (long k = 0; k < g->nG; k++) w[k] *= scale;
    o_dst2(x, w, g->nx - 1, g->ny - 1, &P->p1, &P->p2);
}

static void o_Ainv(const oprob* P, double* x, const double* b, int lev)
{ o_dst_inv(P, x, b, P->divAinv + (long)(lev - 1) * P->g.nG, P->lap_scale); }
static void o_A(const oprob* P, double* x, const double* b, int lev)
{ o_dst_inv(P, x, b, P->divA + (long)(lev - 1) * P->g.nG, P->lap_scale); }
static void o_Dinv(const oprob* P, double* x, const double* b)
{ o_dst_inv(P, x, b, P->Lam, P->lap_scale); }

/* vort2flux! -- curl.jl:133-165.  Mutates the coarse levels of G. */
static void o_vort2flux(oprob* P, double* psi, double* q, double* G, int ngrids)
{
    const og_grid* g = &P->g;
    const long nG = g->nG, nq = g->nq;
    double* psibc = P->Gbc;
    double* Gwork = P->G3;
    for (int lev = 2; lev <= ngrids; lev++) o_coarsify(g, G + (lev - 1) * nG, G + (lev - 2) * nG);
    for (long k = 0; k < nG * g->mg; k++) psi[k] *= 0.0;
    for (int k = 0; k < g->nbc; k++) psibc[k] *= 0.0;
    o_Dinv(P, psi + (ngrids - 1) * nG, G + (ngrids - 1) * nG);
    o_curl_mg(g, q + (ngrids - 1) * nq, psi + (ngrids - 1) * nG, psibc);
    for (int lev = ngrids - 1; lev >= 1; lev--) {
        memcpy(Gwork, G + (lev - 1) * nG, sizeof(double) * (size_t)nG);
        o_get_bc(g, psibc, psi + lev * nG);
        o_apply_bc(g, Gwork, psibc, 1.0);
        o_Dinv(P, psi + (lev - 1) * nG, Gwork);
        o_curl_mg(g, q + (lev - 1) * nq, psi + (lev - 1) * nG, psibc);
    }
}

/* avg_flux -- nonlinear.jl:17-38 (IBProblem method) */
static void o_avg_flux(const og_grid* g, double* Q, const double* q, const double* q0)
{
    const int nx = g->nx, ny = g->ny, nu = g->nu;
    for (long k = 0; k < g->nq; k++) Q[k] *= 0.0;
    for (int j = 2; j <= ny; j++)
        for (int i = 1; i <= nx + 1; i++)
            QX(Q, i, j) = ((QX(q, i, j) + QX(q0, i, j)) + QX(q, i, j - 1)) + QX(q0, i, j - 1);
    for (int j = 1; j <= ny + 1; j++)
        for (int i = 2; i <= nx; i++)
            QY(Q, i, j) = ((QY(q, i, j) + QY(q0, i, j)) + QY(q, i - 1, j)) + QY(q0, i - 1, j);
}

/* direct_product_loops! -- nonlinear.jl:73-114 (nl_avg! :2-10) */
static void o_direct_product_loops(const og_grid* g, double* fq, const double* Q, const double* G, const double* Gbc)
{
    const int nx = g->nx, ny = g->ny, nu = g->nu;
    const int T = g->TOP, B = g->BOT, L = g->LEFT, R = g->RIGHT;
    for (int j = 2; j <= ny - 1; j++)
        for (int i = 2; i <= nx; i++)
            QX(fq, i, j) = QY(Q, i, j + 1) * GM(G, i - 1, j) + QY(Q, i, j) * GM(G, i - 1, j - 1);
    for (int i = 2; i <= nx; i++)
        QX(fq, i, 1) = QY(Q, i, 2) * GM(G, i - 1, 1) + QY(Q, i, 1) * BC(Gbc, B + i);
    for (int i = 2; i <= nx; i++)
        QX(fq, i, ny) = QY(Q, i, ny) * GM(G, i - 1, ny - 1) + QY(Q, i, ny + 1) * BC(Gbc, T + i);
    for (int j = 2; j <= ny; j++)
        for (int i = 2; i <= nx - 1; i++)
            QY(fq, i, j) = (QX(Q, i + 1, j) * GM(G, i, j - 1) + QX(Q, i, j) * GM(G, i - 1, j - 1)) * -1.0;
    for (int j = 2; j <= ny; j++)
        QY(fq, 1, j) = (QX(Q, 2, j) * GM(G, 1, j - 1) + QX(Q, 1, j) * BC(Gbc, L + j)) * -1.0;
    for (int j = 2; j <= ny; j++)
        QY(fq, nx, j) = (QX(Q, nx, j) * GM(G, nx - 1, j - 1) + QX(Q, nx + 1, j) * BC(Gbc, R + j)) * -1.0;
}

/* avg_flux -- nonlinear.jl:45-65 (LinearizedIBProblem method: no base flux) */
static void o_avg_flux_lin(const og_grid* g, double* Q, const double* q)
{
    const int nx = g->nx, ny = g->ny, nu = g->nu;
    for (long k = 0; k < g->nq; k++) Q[k] *= 0.0;
    for (int j = 2; j <= ny; j++)
        for (int i = 1; i <= nx + 1; i++) QX(Q, i, j) = QX(q, i, j) + QX(q, i, j - 1);
    for (int j = 1; j <= ny + 1; j++)
        for (int i = 2; i <= nx; i++) QY(Q, i, j) = QY(q, i, j) + QY(q, i - 1, j);
}

/* nonlinear! -- nonlinear.jl:168-201.  Scales Gbc by 0.25 in place.  For a
 * LinearizedIBProblem the direct product is Q x ΓB + QB x Γ
 * (direct_product!, nonlinear.jl:144-163), both with the same Γbc. */
static void o_nonlinear(oprob* P, double* nonlin, const double* q, const double* q0, const double* G,
                        double* Gbc, int lev)
{
    const og_grid* g = &P->g;
    double* fq = P->q3;
    for (long k = 0; k < g->nq; k++) fq[k] *= 0.0;
    for (int k = 0; k < g->nbc; k++) Gbc[k] *= 0.25;
    double* Q = P->q2;
    if (P->lin) o_avg_flux_lin(g, Q, q);
    else o_avg_flux(g, Q, q, q0);
    for (long k = 0; k < g->nq; k++) fq[k] *= 0.0;
    if (P->lin) {
        double* fq_tmp = P->q4;
        for (long k = 0; k < g->nq; k++) fq_tmp[k] *= 0.0;
        o_direct_product_loops(g, fq, Q, P->GB + (long)(lev - 1) * g->nG, Gbc);
        o_direct_product_loops(g, fq_tmp, P->QB + (long)(lev - 1) * g->nq, G, Gbc);
        for (long k = 0; k < g->nq; k++) fq[k] += fq_tmp[k];
    } else {
        o_direct_product_loops(g, fq, Q, G, Gbc);
    }
    o_rot(g, nonlin, fq);
    double hc = g->h * (double)(1L << (lev - 1));
    double s = 1.0 / (4.0 * (hc * hc));
    for (long k = 0; k < g->nG; k++) nonlin[k] *= s;
}

/* ------------------------------------------------------------------------- */
/* interface coupling (src/interface-coupling/interface-coupling.jl)          */
/* ------------------------------------------------------------------------- */

/* δh -- interface-coupling.jl:6-62 (Yang-3 smoothed delta times h) */
double oracle_delta_h(double rf, double rb, double dr)
{
    double r = fabs(rf - rb);
    double r1 = r / dr, r2 = r1 * r1, r3 = r2 * r1, r4 = r3 * r1;
    double del_h;
    if (r1 <= 1.0) {
        double a5 = asin((1.0 / 2.0) * sqrt(3.0) * (2.0 * r1 - 1.0));
        double a8 = sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
        del_h = 4.166666667e-2 * r4 + (-0.1388888889 + 3.472222222e-2 * a8) * r3 +
                (-7.121664902e-2 - 5.208333333e-2 * a8 + 0.2405626122 * a5) * r2 +
                (-.2405626122 * a5 - 0.3792313933 + 0.1012731481 * a8) * r1 + 8.0187537413e-2 * a5 -
                4.195601852e-2 * a8 + 0.6485698427;
    } else if (r1 <= 2.0) {
        double a6 = asin((1.0 / 2.0) * sqrt(3.0) * (-3.0 + 2.0 * r1));
        double a9 = sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
        del_h = -6.250000000e-2 * r4 + (0.4861111111 - 1.736111111e-2 * a9) * r3 +
                (-1.143175026 + 7.812500000e-2 * a9 - 0.1202813061 * a6) * r2 +
                (0.8751991178 + 0.3608439183 * a6 - 0.1548032407 * a9) * r1 - 0.2806563809 * a6 +
                8.22848104e-3 + 0.1150173611 * a9;
    } else if (r1 <= 3.0) {
        double a1 = asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * sqrt(3.0));
        double a7 = sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
        del_h = 2.083333333e-2 * r4 + (3.472222222e-3 * a7 - 0.2638888889) * r3 +
                (1.214391675 - 2.604166667e-2 * a7 + 2.405626122e-2 * a1) * r2 +
                (-0.1202813061 * a1 - 2.449273192 + 7.262731481e-2 * a7) * r1 + 0.1523563211 * a1 +
                1.843201677 - 7.306134259e-2 * a7;
    } else {
        del_h = 0.0;
    }
    return del_h;
}

/* setup_reg -- interface-coupling.jl:72-96 */
static void o_setup_reg(oprob* P)
{
    const og_grid* g = &P->g;
    const int nbt = P->nbt;
    const double h = g->h;
    for (int k = 0; k < nbt; k++) {
        P->body_idx[k] = (int)floor((P->xb[k] + g->offx) / h);
        P->body_idx[k + nbt] = (int)floor((P->xb[k + nbt] + g->offy) / h);
    }
    for (int k = 0; k < nbt; k++) {
        double xk = P->xb[k], yk = P->xb[k + nbt];
        int bx = P->body_idx[k], by = P->body_idx[k + nbt];
        double* w1 = P->weight + (long)(2 * k) * WIN * WIN;
        double* w2 = P->weight + (long)(2 * k + 1) * WIN * WIN;
        for (int m = 0; m < WIN; m++) {
            double y = h * (double)(by - 1 + (m - SUPP)) - g->offy;
            for (int l = 0; l < WIN; l++) {
                double x = h * (double)(bx - 1 + (l - SUPP)) - g->offx;
                w1[l + WIN * m] = oracle_delta_h(x, xk, h) * oracle_delta_h(y + h / 2, yk, h);
                w2[l + WIN * m] = oracle_delta_h(x + h / 2, xk, h) * oracle_delta_h(y, yk, h);
            }
        }
    }
}

/* reg! (E') -- interface-coupling.jl:99-119 */
static void o_reg(const oprob* P, double* q, const double* fb)
{
    const og_grid* g = &P->g;
    const int nx = g->nx, nu = g->nu, nbt = P->nbt;
    (void)nx;
    for (long k = 0; k < g->nq; k++) q[k] = 0.0;
    for (int k = 0; k < nbt; k++) {
        int bx = P->body_idx[k], by = P->body_idx[k + nbt];
        const double* w1 = P->weight + (long)(2 * k) * WIN * WIN;
        const double* w2 = P->weight + (long)(2 * k + 1) * WIN * WIN;
        double fx = fb[k], fy = fb[k + nbt];
        for (int m = 0; m < WIN; m++)
            for (int l = 0; l < WIN; l++) {
                int i = bx + l - SUPP, j = by + m - SUPP;
                QX(q, i, j) = QX(q, i, j) + w1[l + WIN * m] * fx;
                QY(q, i, j) = QY(q, i, j) + w2[l + WIN * m] * fy;
            }
    }
}

/* regT! (E) -- interface-coupling.jl:122-133 */
static void o_regT(const oprob* P, double* fb, const double* q)
{
    const og_grid* g = &P->g;
    const int nx = g->nx, nu = g->nu, nbt = P->nbt;
    (void)nx;
    for (int k = 0; k < 2 * nbt; k++) fb[k] = 0.0;
    for (int k = 0; k < nbt; k++) {
        int bx = P->body_idx[k], by = P->body_idx[k + nbt];
        const double* w1 = P->weight + (long)(2 * k) * WIN * WIN;
        const double* w2 = P->weight + (long)(2 * k + 1) * WIN * WIN;
        double s1 = 0.0, s2 = 0.0;
        for (int m = 0; m < WIN; m++)
            for (int l = 0; l < WIN; l++) {
                int i = bx + l - SUPP, j = by + m - SUPP;
                s1 += w1[l + WIN * m] * QX(q, i, j);
                s2 += w2[l + WIN * m] * QY(q, i, j);
            }
        fb[k] += s1;
        fb[k + nbt] += s2;
    }
}

/* B_times! -- laplacian.jl:143-169 */
static void o_B_times(oprob* P, double* x, const double* z)
{
    const og_grid* g = &P->g;
    const long nG = g->nG;
    double* G = P->bt_G;
    for (long k = 0; k < nG * g->mg; k++) G[k] *= 0.0;
    /* Γ[:,1] = Ainv * ((E*C)' * z) = Ainv(rot(reg(z))) */
    o_reg(P, P->q4, z);
    o_rot(g, P->Gwork, P->q4);
    o_Ainv(P, G, P->Gwork, 1);
    o_vort2flux(P, P->bt_psi, P->bt_q, G, g->mg);
    o_regT(P, x, P->bt_q);
}

/* in-place inverse of a dense column-major n x n matrix by LU with partial
 * pivoting (the algorithm of LAPACK getrf/getri, which inv() calls). */
static int o_invert(double* A, int n)
{
    int* piv = (int*)malloc(sizeof(int) * (size_t)n);
    double* LU = (double*)malloc(sizeof(double) * (size_t)n * n);
    memcpy(LU, A, sizeof(double) * (size_t)n * n);
#define LUA(i, j) LU[(i) + (long)n * (j)]
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(LUA(k, k));
        for (int i = k + 1; i < n; i++) if (fabs(LUA(i, k)) > best) { best = fabs(LUA(i, k)); p = i; }
        piv[k] = p;
        if (best == 0.0) { free(piv); free(LU); return -1; }
        if (p != k) for (int j = 0; j < n; j++) { double t = LUA(k, j); LUA(k, j) = LUA(p, j); LUA(p, j) = t; }
        double inv = 1.0 / LUA(k, k);
        for (int i = k + 1; i < n; i++) LUA(i, k) *= inv;
#pragma omp parallel for num_threads(o_nthreads) schedule(static)
        for (int j = k + 1; j < n; j++) {
            double akj = LUA(k, j);
            if (akj != 0.0) for (int i = k + 1; i < n; i++) LUA(i, j) -= LUA(i, k) * akj;
        }
    }
    /* solve A X = I column by column: P A = L U */
#pragma omp parallel for num_threads(o_nthreads) schedule(static)
    for (int c = 0; c < n; c++) {
        double* x = A + (long)n * c;
        for (int i = 0; i < n; i++) x[i] = 0.0;
        x[c] = 1.0;
        for (int k = 0; k < n; k++) if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
        for (int j = 0; j < n; j++) { double xj = x[j]; if (xj != 0.0) for (int i = j + 1; i < n; i++) x[i] -= LUA(i, j) * xj; }
        for (int j = n - 1; j >= 0; j--) { x[j] /= LUA(j, j); double xj = x[j]; for (int i = 0; i < j; i++) x[i] -= LUA(i, j) * xj; }
    }
#undef LUA
    free(piv); free(LU);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* public API (ctypes)                                                        */
/* ------------------------------------------------------------------------- */
typedef struct {
    double *q, *q0, *G, *psi, *nl1, *nl2, *Fb, *fb, *CD, *CL;
} ostate;

void oracle_problem_free(oprob* P);

/* IBProblem(grid, bodies, dt, Re) -- problem-types.jl:23-41 with IBModel
 * (models.jl:124-158), IBMatrices (models.jl:33-52) and get_AB
 * (laplacian.jl:83-89).  When binv == NULL, Binv is formed by get_Binv
 * (laplacian.jl:98-119); otherwise it is copied from binv. */
oprob* oracle_problem_create(double h, double x0, double x1, double y0, double y1, int mg,
                             double Re, double dt, int nbodies, const int* nb, const double* xb,
                             const double* ds, int motion, const double* binv, int* status)
{
    oprob* P = (oprob*)calloc(1, sizeof(oprob));
    og_grid* g = &P->g;
    og_make_grid(g, h, x0, x1, y0, y1, mg);
    const int nx = g->nx, ny = g->ny;
    P->Re = Re; P->dt = dt;
    P->nbeta = 2; P->beta[0] = 1.5; P->beta[1] = -0.5; /* AB2 -- models.jl:70-72 */
    P->nbodies = nbodies;
    P->nb = (int*)malloc(sizeof(int) * (size_t)(nbodies > 0 ? nbodies : 1));
    int nbt = 0;
    for (int b = 0; b < nbodies; b++) { P->nb[b] = nb[b]; nbt += nb[b]; }
    P->nbt = nbt;
    P->xb = (double*)malloc(sizeof(double) * 2 * (size_t)(nbt > 0 ? nbt : 1));
    P->ds = (double*)malloc(sizeof(double) * (size_t)(nbt > 0 ? nbt : 1));
    memcpy(P->xb, xb, sizeof(double) * 2 * (size_t)nbt);
    memcpy(P->ds, ds, sizeof(double) * (size_t)nbt);
    P->motion = motion;
    P->x0 = (double*)malloc(sizeof(double) * 2 * (size_t)(nbt > 0 ? nbt : 1));
    P->ub = (double*)calloc(2 * (size_t)(nbt > 0 ? nbt : 1), sizeof(double));
    memcpy(P->x0, xb, sizeof(double) * 2 * (size_t)nbt); /* x0 = copy(xb), sample-bodies.jl:32 */
    const long nG = g->nG, nq = g->nq;
    P->Lam = (double*)malloc(sizeof(double) * nG);
    o_lap_eigs(g, P->Lam);
    P->lap_scale = 1.0 / (double)(4 * nx * ny);
    P->divA = (double*)malloc(sizeof(double) * nG * mg);
    P->divAinv = (double*)malloc(sizeof(double) * nG * mg);
    for (int lev = 1; lev <= mg; lev++) {
        double hc = g->h * (double)(1L << (lev - 1));
        double c = (2.0 * Re) * (hc * hc);
        for (long k = 0; k < nG; k++) {
            double lt = 1.0 - (P->Lam[k] * dt) / c;                    /* get_A    laplacian.jl:74 */
            P->divA[(lev - 1) * nG + k] = 1.0 / lt;                      /* 1 ./ Λ̃   laplacian.jl:75 */
            P->divAinv[(lev - 1) * nG + k] = 1.0 + (P->Lam[k] * dt) / c; /* get_Ainv laplacian.jl:69 */
        }
    }
    odst_plan_init(&P->p1, nx - 1);
    odst_plan_init(&P->p2, ny - 1);
    /* WorkingMemory -- models.jl:91-105 */
    P->q1 = (double*)calloc(nq * mg, sizeof(double));
    P->G1 = (double*)calloc(nG * mg, sizeof(double));
    P->q2 = (double*)calloc(nq, sizeof(double));
    P->G2 = (double*)calloc(nG, sizeof(double));
    P->q3 = (double*)calloc(nq, sizeof(double));
    P->q4 = (double*)calloc(nq, sizeof(double));
    P->q5 = (double*)calloc(nq, sizeof(double));
    P->q6 = (double*)calloc(nq, sizeof(double));
    P->G3 = (double*)calloc(nG, sizeof(double));
    P->Gbc = (double*)calloc(g->nbc, sizeof(double));
    P->rhsbc = (double*)calloc(nG, sizeof(double));
    P->Gwork = (double*)calloc(nG, sizeof(double));
    P->wdst = (double*)calloc(nG, sizeof(double));
    P->bt_G = (double*)calloc(nG * mg, sizeof(double));
    P->bt_psi = (double*)calloc(nG * mg, sizeof(double));
    P->bt_q = (double*)calloc(nq * mg, sizeof(double));
    /* E -- setup_reg */
    P->body_idx = (int*)malloc(sizeof(int) * 2 * (size_t)(nbt > 0 ? nbt : 1));
    P->weight = (double*)calloc((size_t)(2 * (nbt > 0 ? nbt : 1)) * WIN * WIN, sizeof(double));
    o_setup_reg(P);
    *status = 0;
    for (int k = 0; k < nbt; k++) {
        int bx = P->body_idx[k], by = P->body_idx[k + nbt];
        if (bx - SUPP < 1 || bx + SUPP > nx || by - SUPP < 1 || by + SUPP > ny) *status = -2;
    }
    if (*status) { oracle_problem_free(P); return NULL; }
    /* moving-grid coordinates -- models.jl:135-153 (xc = yc = 0) */
    if (motion == 1) {
        P->XX = (double*)malloc(sizeof(double) * (size_t)nx * (ny + 1) * mg);
        P->YY = (double*)malloc(sizeof(double) * (size_t)ny * (nx + 1) * mg);
        for (int lev = 1; lev <= mg; lev++) {
            double hc = g->h * (double)(1L << (lev - 1));
            for (int j = 1; j <= ny; j++) {
                double y = (((double)j - 0.5) - (double)ny / 2.0) * hc + ((double)ny / 2.0) * g->h - g->offy;
                for (int i = 1; i <= nx + 1; i++) P->YY[(long)(lev - 1) * g->nu + (i - 1) + (long)(nx + 1) * (j - 1)] = y - 0.0;
            }
            for (int j = 1; j <= ny + 1; j++)
                for (int i = 1; i <= nx; i++) {
                    double x = (((double)i - 0.5) - (double)nx / 2.0) * hc + ((double)nx / 2.0) * g->h - g->offx;
                    P->XX[(long)(lev - 1) * g->nv + (i - 1) + (long)nx * (j - 1)] = x - 0.0;
                }
        }
    }
    const int nf = 2 * nbt;
    P->Binv = (double*)calloc((size_t)(nf > 0 ? nf : 1) * (nf > 0 ? nf : 1), sizeof(double));
    if (motion == 2) {
        /* get_Binv for MotionFunction (laplacian.jl:123-141) forms no matrix:
         * prob.Binv is a CG solve against B_times!, done per step */
    } else if (binv) {
        memcpy(P->Binv, binv, sizeof(double) * (size_t)nf * nf);
    } else {
        double* e = (double*)calloc(nf, sizeof(double));
        for (int j = 0; j < nf; j++) {
            for (int k = 0; k < nf; k++) e[k] *= 0.0;
            e[j] = 1.0;
            o_B_times(P, P->Binv + (long)nf * j, e);
        }
        free(e);
        if (nf > 0 && o_invert(P->Binv, nf) != 0) { *status = -3; oracle_problem_free(P); return NULL; }
    }
    return P;
}

void oracle_problem_free(oprob* P)
{
    if (!P) return;
    odst_plan_free(&P->p1); odst_plan_free(&P->p2);
    free(P->nb); free(P->xb); free(P->ds); free(P->x0); free(P->ub); free(P->Lam); free(P->divA); free(P->divAinv);
    free(P->body_idx); free(P->weight); free(P->Binv); free(P->XX); free(P->YY);
    free(P->q1); free(P->G1); free(P->q2); free(P->G2); free(P->q3); free(P->q4); free(P->q5); free(P->q6);
    free(P->G3); free(P->Gbc); free(P->rhsbc); free(P->Gwork); free(P->wdst);
    free(P->bt_G); free(P->bt_psi); free(P->bt_q);
    free(P->QB); free(P->GB);
    free(P);
}

void oracle_problem_sizes(const oprob* P, int* out /* nx ny nG nq mg nbt nbc */, double* dout /* h offx offy len */)
{
    out[0] = P->g.nx; out[1] = P->g.ny; out[2] = P->g.nG; out[3] = P->g.nq; out[4] = P->g.mg;
    out[5] = P->nbt; out[6] = P->g.nbc;
    dout[0] = P->g.h; dout[1] = P->g.offx; dout[2] = P->g.offy; dout[3] = P->g.len;
}

void oracle_get_binv(const oprob* P, double* out) { memcpy(out, P->Binv, sizeof(double) * 4 * (size_t)P->nbt * P->nbt); }
void oracle_get_body_idx(const oprob* P, int* out) { memcpy(out, P->body_idx, sizeof(int) * 2 * (size_t)P->nbt); }
void oracle_get_weights(const oprob* P, double* out) { memcpy(out, P->weight, sizeof(double) * 2 * (size_t)P->nbt * WIN * WIN); }
void oracle_get_lam(const oprob* P, double* out) { memcpy(out, P->Lam, sizeof(double) * (size_t)P->g.nG); }

/* base_flux! for InertialMotion -- timestepping.jl:252-270 (used by IBState) */
static void o_base_flux_inertial(const oprob* P, double* q0, double Ux, double Uy, double alpha)
{
    const og_grid* g = &P->g;
    const int nu = g->nu;
    for (int lev = 1; lev <= g->mg; lev++) {
        double hc = g->h * (double)(1L << (lev - 1));
        double vx = (Ux * cos(alpha) - Uy * sin(alpha)) * hc;
        double vy = (Ux * sin(alpha) + Uy * cos(alpha)) * hc;
        double* col = q0 + (long)(lev - 1) * g->nq;
        for (int k = 0; k < nu; k++) col[k] = vx;
        for (long k = nu; k < g->nq; k++) col[k] = vy;
    }
}

/* base_flux! for MovingGrid -- timestepping.jl:273-318 */
static void o_base_flux_moving(const oprob* P, double* q0, double U, double V, double th, double thdot,
                               double Uxf, double Uyf, double alf)
{
    const og_grid* g = &P->g;
    const int nu = g->nu;
    const long nq = g->nq;
    double Om = -thdot;
    double al = -th;
    double Ux0 = U * cos(al) - V * sin(al);
    double Uy0 = U * sin(al) + V * cos(al);
    Ux0 += Uxf * cos(alf) - Uyf * sin(alf);
    Uy0 += Uxf * sin(alf) + Uyf * cos(alf);
    for (long k = 0; k < nq * g->mg; k++) q0[k] *= 0.0;
    for (int lev = 1; lev <= g->mg; lev++) {
        double hc = g->h * pow(2.0, (double)lev - 1.0);
        double* col = q0 + (long)(lev - 1) * nq;
        const double* YY = P->YY + (long)(lev - 1) * nu;
        const double* XX = P->XX + (long)(lev - 1) * g->nv;
        double sx = -hc * Om, sy = hc * Om;
        for (int k = 0; k < nu; k++) { col[k] = YY[k]; col[k] *= sx; }
        for (long k = nu; k < nq; k++) { col[k] = XX[k - nu]; col[k] *= sy; }
        for (int k = 0; k < nu; k++) col[k] += hc * Ux0;
        for (long k = nu; k < nq; k++) col[k] += hc * Uy0;
    }
}

/* scalars[] = {Ux, Uy, inclination, U, V, theta, theta_dot} evaluated at t.
 * A LinearizedIBProblem with static bodies has q0 = 0 (timestepping.jl:238-250). */
void oracle_base_flux(const oprob* P, ostate* S, const double* sc)
{
    if (P->lin) {
        for (long k = 0; k < P->g.nq * P->g.mg; k++) S->q0[k] = 0.0;
        return;
    }
    if (P->motion == 1) o_base_flux_moving(P, S->q0, sc[3], sc[4], sc[5], sc[6], sc[0], sc[1], sc[2]);
    else o_base_flux_inertial(P, S->q0, sc[0], sc[1], sc[2]);
}

/* LinearizedIBProblem(base_state, base_prob, dt) -- problem-types.jl:66-86:
 * QB[:, lev] = avg_flux(base_state, base_prob; lev) (the IBProblem method,
 * with the base q0), ΓB = base_state.Γ.  Static bodies only (returns -1
 * otherwise). */
int oracle_problem_linearize(oprob* P, const double* qB, const double* q0B, const double* GB)
{
    if (P->motion != 0) return -1;
    const og_grid* g = &P->g;
    const long nq = g->nq, nG = g->nG;
    if (!P->QB) P->QB = (double*)calloc((size_t)(nq * g->mg), sizeof(double));
    if (!P->GB) P->GB = (double*)calloc((size_t)(nG * g->mg), sizeof(double));
    for (int lev = 1; lev <= g->mg; lev++)
        o_avg_flux(g, P->QB + (long)(lev - 1) * nq, qB + (long)(lev - 1) * nq, q0B + (long)(lev - 1) * nq);
    memcpy(P->GB, GB, sizeof(double) * (size_t)(nG * g->mg));
    P->lin = 1;
    return 0;
}

/* get_trial_state! -- timestepping.jl:86-129 */
static void o_get_trial_state(oprob* P, double* qs, double* Gs, ostate* S)
{
    const og_grid* g = &P->g;
    const long nG = g->nG, nq = g->nq;
    const double dt = P->dt;
    double* rhsbc = P->rhsbc;
    double* rhs = P->G2;
    double* bc = P->Gbc;
    for (int lev = g->mg; lev >= 1; lev--) {
        for (int k = 0; k < g->nbc; k++) bc[k] *= 0.0;
        for (long k = 0; k < nG; k++) rhsbc[k] *= 0.0;
        double hc = g->h * (double)(1L << (lev - 1));
        if (lev < g->mg) {
            o_get_bc(g, bc, S->G + (long)lev * nG);
            double fac = 0.25 * dt / (P->Re * (hc * hc));
            o_apply_bc(g, rhsbc, bc, fac);
        }
        o_nonlinear(P, S->nl1 + (lev - 1) * nG, S->q + (lev - 1) * nq, S->q0 + (lev - 1) * nq,
                    S->G + (lev - 1) * nG, bc, lev);
        o_A(P, rhs, S->G + (lev - 1) * nG, lev);
        double c1 = P->beta[0] * dt, c2 = P->beta[1] * dt;
        const double* n1 = S->nl1 + (lev - 1) * nG;
        const double* n2 = S->nl2 + (lev - 1) * nG;
        for (long k = 0; k < nG; k++) rhs[k] += c1 * n1[k];
        for (long k = 0; k < nG; k++) rhs[k] += c2 * n2[k];
        for (long k = 0; k < nG; k++) rhs[k] += rhsbc[k];
        o_Ainv(P, Gs + (lev - 1) * nG, rhs, lev);
    }
    memcpy(S->nl2, S->nl1, sizeof(double) * nG * g->mg);
    o_vort2flux(P, S->psi, qs, Gs, g->mg);
}

/* boundary_forces! (Static / MovingGrid) -- timestepping.jl:154-163 */
static void o_boundary_forces(oprob* P, double* Fb, const double* qs1, const double* q01)
{
    const og_grid* g = &P->g;
    const int nf = 2 * P->nbt;
    double* Q = P->q2;
    for (long k = 0; k < g->nq; k++) Q[k] = qs1[k] + q01[k];
    double* tmp = (double*)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    o_regT(P, tmp, Q);
    for (int i = 0; i < nf; i++) {
        double s = 0.0;
        for (int j = 0; j < nf; j++) s += P->Binv[i + (long)nf * j] * tmp[j];
        Fb[i] = s;
    }
    free(tmp);
}

/* move_body!(::Type{MotionFunction}, body, t) -- move-body-utils.jl:123-133,
 * with get_transformation -- :73-97.  kin = {x, y, θ, ẋ, ẏ, θ̇} per body, the
 * closures motion.xc[1:3](t), motion.uc[1:3](t):
 *   xb = pos + Rx x0,  Rx = [cθ -sθ; sθ cθ]
 *   ub = vel + Ru x0,  Ru = θ̇ [-sθ -cθ; cθ -sθ]                               */
static void o_move_bodies(oprob* P, const double* kin)
{
    const int nbt = P->nbt;
    int tally = 0;
    for (int b = 0; b < P->nbodies; b++) {
        const double* k = kin + 6L * b;
        const double c = cos(k[2]), s = sin(k[2]);
        const double Rx11 = c, Rx12 = -s, Rx21 = s, Rx22 = c;
        const double Ru11 = k[5] * -s, Ru12 = k[5] * -c, Ru21 = k[5] * c, Ru22 = k[5] * -s;
        for (int i = 0; i < P->nb[b]; i++) {
            const long ix = tally + i, iy = (long)nbt + tally + i;
            const double x0 = P->x0[ix], y0 = P->x0[iy];
            P->xb[ix] = k[0] + (Rx11 * x0 + Rx12 * y0);
            P->xb[iy] = k[1] + (Rx21 * x0 + Rx22 * y0);
            P->ub[ix] = k[3] + (Ru11 * x0 + Ru12 * y0);
            P->ub[iy] = k[4] + (Ru21 * x0 + Ru22 * y0);
        }
        tally += P->nb[b];
    }
}

/* update_bodies!(::Type{MotionFunction}, state, prob, t) --
 * timestepping.jl:386-398: move every body, then E = setup_reg(grid, bodies)
 * (the new Binv is the CG map, nothing to form).  Returns -2 when a body
 * point comes within 6 cells of the domain edge (the 13 x 13 window of E
 * would leave the grid; Julia raises a BoundsError there). */
static int o_update_bodies(oprob* P, const double* kin)
{
    o_move_bodies(P, kin);
    o_setup_reg(P);
    const int nx = P->g.nx, ny = P->g.ny;
    for (int k = 0; k < P->nbt; k++) {
        int bx = P->body_idx[k], by = P->body_idx[k + P->nbt];
        if (bx - SUPP < 1 || bx + SUPP > nx || by - SUPP < 1 || by + SUPP > ny) return -2;
    }
    return 0;
}

static double o_dot(const double* a, const double* b, int n)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}

/* cg!(x, B, b; maxiter, reltol) of IterativeSolvers.jl 0.9.1 (the reference's
 * Manifest.toml:266-270; the package is not in the reference tree), as called
 * by prob.Binv of get_Binv (laplacian.jl:137-138) with B = B_times!:
 * cg_iterator! with initially_zero = false (x is the initial guess; the
 * mutating LinearMap hands cg! the output array, i.e. the state's F̃b of the
 * previous step), tolerance = max(reltol |b|, abstol = 0), then the loop
 *   β = |r|² / |r_prev|²  (|r_prev| = 1 at the start, u = 0),  u = r + β u,
 *   c = B u,  α = |r|² / (u·c),  x += α u,  r -= α c
 * until |r| <= tolerance or maxiter products.  |·| and · are plain sums here
 * (Julia uses BLAS nrm2 and dot: equal up to rounding). */
static void o_cg(oprob* P, double* x, const double* b, double reltol, int maxiter)
{
    const int n = 2 * P->nbt;
    double* u = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* r = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* c = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    memcpy(r, b, sizeof(double) * (size_t)n);
    o_B_times(P, c, x);
    for (int k = 0; k < n; k++) r[k] -= c[k];
    double res = sqrt(o_dot(r, r, n));
    const double tol = fmax(reltol * sqrt(o_dot(b, b, n)), 0.0);
    double prev = 1.0;
    int it = 0;
    while (!(it >= maxiter || res <= tol)) {
        const double beta = (res * res) / (prev * prev);
        for (int k = 0; k < n; k++) u[k] = r[k] + beta * u[k];
        o_B_times(P, c, u);
        const double alpha = (res * res) / o_dot(u, c, n);
        for (int k = 0; k < n; k++) x[k] += alpha * u[k];
        for (int k = 0; k < n; k++) r[k] -= alpha * c[k];
        prev = res;
        res = sqrt(o_dot(r, r, n));
        it++;
    }
    P->cg_iters = it;
    P->cg_res = res;
    P->cg_tol = tol;
    free(u); free(r); free(c);
}

/* boundary_forces! for V <: Motion -- timestepping.jl:176-186:
 * F̃work = E (qs + q0) - ub h (ub stacked [ub; vb], get_ub move-body-utils.jl:35-48),
 * F̃b = Binv F̃work with Binv the CG map (laplacian.jl:137-138: maxiter 5000,
 * reltol 1e-12). */
static void o_boundary_forces_mf(oprob* P, double* Fb, const double* qs1, const double* q01)
{
    const og_grid* g = &P->g;
    const int nf = 2 * P->nbt;
    const double h = g->h;
    double* Q = P->q2;
    for (long k = 0; k < g->nq; k++) Q[k] = qs1[k] + q01[k];
    double* Fw = (double*)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
    o_regT(P, Fw, Q);
    for (int k = 0; k < nf; k++) Fw[k] -= P->ub[k] * h;
    o_cg(P, Fb, Fw, 1e-12, 5000);
    free(Fw);
}

/* update_stress! -- timestepping.jl:330-354 */
static void o_update_stress(const oprob* P, ostate* S)
{
    const double h = P->g.h, dt = P->dt;
    const int nbt = P->nbt;
    int tally = 0;
    for (int b = 0; b < P->nbodies; b++) {
        int n = P->nb[b];
        double* fb = S->fb + 2L * tally; /* body b: n x 2 col-major */
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < n; k++) {
            double ds = P->ds[tally + k];
            fb[k] = (S->Fb[tally + k] * (h / dt)) / ds;
            fb[n + k] = (S->Fb[nbt + tally + k] * (h / dt)) / ds;
        }
        for (int k = 0; k < n; k++) sx += P->ds[tally + k] * fb[k];
        for (int k = 0; k < n; k++) sy += P->ds[tally + k] * fb[n + k];
        S->CD[b] = 2.0 * sx;
        S->CL[b] = 2.0 * sy;
        tally += n;
    }
}

/* project_circ! -- timestepping.jl:204-223 */
static void o_project_circ(oprob* P, double* Gs, ostate* S)
{
    const og_grid* g = &P->g;
    const long nG = g->nG;
    double* Gwork = P->G2;
    memcpy(S->G, Gs, sizeof(double) * nG * g->mg);
    o_reg(P, P->q4, S->Fb);
    o_rot(g, Gs, P->q4);
    o_Ainv(P, Gwork, Gs, 1);
    for (long k = 0; k < nG; k++) S->G[k] -= Gwork[k];
}

/* advance!(state, prob, t) -- timestepping.jl:35-67.  sc: closures at t
 * (oracle_base_flux's 7 scalars; for MotionFunction followed by the 6
 * kinematic values {x, y, θ, ẋ, ẏ, θ̇} of every body).  Returns 0, or -2 when
 * a moving body leaves the grid. */
int oracle_advance(oprob* P, ostate* S, const double* sc)
{
    if (P->motion == 2) { /* update_bodies!, timestepping.jl:39 */
        int rc = o_update_bodies(P, sc + 7);
        if (rc) return rc;
    }
    if (P->motion == 1) oracle_base_flux(P, S, sc); /* timestepping.jl:41-43 */
    double* qs = P->q1;
    double* Gs = P->G1;
    o_get_trial_state(P, qs, Gs, S);
    if (P->motion == 2) o_boundary_forces_mf(P, S->Fb, qs, S->q0);
    else o_boundary_forces(P, S->Fb, qs, S->q0);
    o_update_stress(P, S);
    o_project_circ(P, Gs, S);
    o_vort2flux(P, S->psi, S->q, S->G, P->g.mg);
    return 0;
}

/* scalars per step of oracle_run: 7, plus 6 per body for MotionFunction */
int oracle_step_stride(const oprob* P) { return 7 + (P->motion == 2 ? 6 * P->nbodies : 0); }

/* run nsteps from times t[0..nsteps-1]; sc: oracle_step_stride scalars per
 * step; CD/CL history (nsteps x nbodies) is recorded when hist != NULL.
 * Returns 0, or the first failing step's status. */
int oracle_run(oprob* P, ostate* S, int nsteps, const double* sc, double* cd_hist, double* cl_hist)
{
    const long stride = oracle_step_stride(P);
    for (int s = 0; s < nsteps; s++) {
        int rc = oracle_advance(P, S, sc + stride * s);
        if (rc) return rc;
        if (cd_hist) for (int b = 0; b < P->nbodies; b++) {
            cd_hist[(long)s * P->nbodies + b] = S->CD[b];
            cl_hist[(long)s * P->nbodies + b] = S->CL[b];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* operator-level entry points for unit tests                                 */
/* ------------------------------------------------------------------------- */
void oracle_op_dst2(const oprob* P, double* y, const double* x) { o_dst2(y, x, P->g.nx - 1, P->g.ny - 1, &P->p1, &P->p2); }

/* raw DST-I of n1 x n2 arrays without a problem */
void oracle_dst2_raw(double* y, const double* x, int n1, int n2)
{
    odst_plan p1, p2;
    odst_plan_init(&p1, n1); odst_plan_init(&p2, n2);
    o_dst2(y, x, n1, n2, &p1, &p2);
    odst_plan_free(&p1); odst_plan_free(&p2);
}

/* kind: 0 = Δinv, 1 = A[lev], 2 = Ainv[lev] */
void oracle_op_lapinv(const oprob* P, double* x, const double* b, int kind, int lev)
{
    if (kind == 0) o_Dinv(P, x, b);
    else if (kind == 1) o_A(P, x, b, lev);
    else o_Ainv(P, x, b, lev);
}
void oracle_op_rot(const oprob* P, double* G, const double* q) { o_rot(&P->g, G, q); }
void oracle_op_curl_sg(const oprob* P, double* q, const double* psi) { o_curl_sg(&P->g, q, psi); }
void oracle_op_curl_mg(const oprob* P, double* q, const double* psi, const double* bc) { o_curl_mg(&P->g, q, psi, bc); }
void oracle_op_get_bc(const oprob* P, double* bc, const double* r) { o_get_bc(&P->g, bc, r); }
void oracle_op_apply_bc(const oprob* P, double* r, const double* bc, double fac) { o_apply_bc(&P->g, r, bc, fac); }
void oracle_op_coarsify(const oprob* P, double* Gc, const double* G) { o_coarsify(&P->g, Gc, G); }
void oracle_op_reg(const oprob* P, double* q, const double* fb) { o_reg(P, q, fb); }
void oracle_op_regT(const oprob* P, double* fb, const double* q) { o_regT(P, fb, q); }
void oracle_op_vort2flux(oprob* P, double* psi, double* q, double* G) { o_vort2flux(P, psi, q, G, P->g.mg); }
/* nonlinear! for level lev with the coarse-level boundary condition taken from
 * G[:, lev+1] (zero on the coarsest level), as get_trial_state! does. */
void oracle_op_nonlinear(oprob* P, double* out, const double* q, const double* q0, const double* G, int lev)
{
    const og_grid* g = &P->g;
    for (int k = 0; k < g->nbc; k++) P->Gbc[k] = 0.0;
    if (lev < g->mg) o_get_bc(g, P->Gbc, G + (long)lev * g->nG);
    o_nonlinear(P, out, q + (long)(lev - 1) * g->nq, q0 + (long)(lev - 1) * g->nq, G + (long)(lev - 1) * g->nG, P->Gbc, lev);
}
void oracle_op_B_times(oprob* P, double* x, const double* z) { o_B_times(P, x, z); }
int oracle_invert(double* A, int n) { return o_invert(A, n); }
/* MotionFunction: current body points and velocities (nbt x 2 col-major
 * each, state.xb and RigidBody.ub), and the last CG solve (iterations,
 * final |r|, tolerance). */
void oracle_get_body_state(const oprob* P, double* xb, double* ub)
{
    if (xb) memcpy(xb, P->xb, sizeof(double) * 2 * (size_t)P->nbt);
    if (ub) memcpy(ub, P->ub, sizeof(double) * 2 * (size_t)P->nbt);
}
void oracle_cg_stats(const oprob* P, int* iters, double* res_tol)
{
    *iters = P->cg_iters;
    res_tol[0] = P->cg_res;
    res_tol[1] = P->cg_tol;
}
/* update_bodies! alone (kin: 6 per body) and one CG solve B x = b from the
 * initial guess in x (reltol, maxiter as given) */
int oracle_op_update_bodies(oprob* P, const double* kin) { return o_update_bodies(P, kin); }
void oracle_op_cg(oprob* P, double* x, const double* b, double reltol, int maxiter) { o_cg(P, x, b, reltol, maxiter); }
