// mixed.h -- argument blocks and launchers of the mixed-precision Laplacian
// inverses (mixed.hip): dense sine-matrix DST-I passes on the MFMA units.
#pragma once
#include "ibpm_dev.h"

namespace ibpm {

// precision of the transforms (IBPM_OPT_PRECISION)
enum MxPrec { MX_FP32 = 0, MX_BF16X3 = 1 };

// residual kinds
enum MxResid {
    MX_RES_TRIAL = 0,  // A Γ - Ainv^{-1} Γ + β1 dt N1 + β2 dt N0 + rhsbc (trial stage, increment form)
    MX_RES_V2F = 1,    // Γ + apply_bc!(get_bc!(ψ coarser)) - L ψ (vort2flux! refinement of ψ)
    MX_RES_COPY = 2,   // the field itself (an operator applied directly)
    MX_RES_TRIAL2 = 3, // (A Γ + β1 dt N1 + β2 dt N0 + rhsbc) - Ainv^{-1} Γs: the trial residual after a sweep
    MX_RES_PROJ2 = 4,  // g - Ainv^{-1} c: the projection's residual after a sweep (G = g, G2 = c)
};

struct MxResidArgs {
    const double* G;    // Γ-like input, entry b at G + b fstride (level lev0 + b)
    const double* G2;   // MX_RES_TRIAL2: Γs after the first sweep; MX_RES_PROJ2: the correction c
    const double* psi;  // MX_RES_V2F: ψ (levels as G)
    const double* N1;   // MX_RES_TRIAL: this step's and the last step's nonlinear terms
    const double* N0;
    long fstride;       // nG
    float* out;         // padded FP32, entry b at out + b ostride
    long ostride;       // P1 P2
    int P1, P2;
    int mode, lev0, mg;
    double c1, c2, dt, Re;
    const double* lscale;  // per-level scale of the FP32 operands (a power of two), indexed by level
    GridDev g;
};
void launch_mx_resid(const MxResidArgs& a, int nbatch, hipStream_t s);
// y += a x (FP64)
void launch_mx_axpy(double* y, const double* x, double a, long n, hipStream_t s);

// GEMM epilogues
enum MxEpi {
    MX_EPI_STORE = 0,  // padded FP32 out
    MX_EPI_SCALE = 1,  // times the FP64 eigenvalue factor, then padded FP32 out
    MX_EPI_ADD = 2,    // FP64 dst = base + acc / lscale (unpadded field layout; base null: 0)
    MX_EPI_SUB = 3,    // FP64 dst = base - acc / lscale
};
enum MxOp { MX_OP_DINV = 0, MX_OP_AINV = 2 };

struct MxGemmArgs {
    const float* in;   // padded FP32 input, entry b at in + b istride
    long istride;
    float* out;        // STORE / SCALE output
    long ostride;
    double* dst;       // ADD / SUB output and its base (may alias), entry b at + b fstride
    const double* base;
    long fstride;
    // sine matrix of the transformed dimension, K x K (K = P1 for x, P2 for y):
    // FP32, or its bf16 high and low parts
    const float* sf;
    const __bf16* sh;
    const __bf16* sl;
    int P1, P2;
    int lev0, fixlev;  // entry b is level lev0 + b (fixlev: every entry at lev0)
    // SCALE: Λ from cx[i] = cos(pi (i+1)/nx), cy[j] = cos(pi (j+1)/ny)
    int op;
    const double* cx;
    const double* cy;
    double dt, Re, scale;
    const double* lscale;
    GridDev g;
    // split-K scratch: MX_SPLIT_MAX x (entries) x P1 P2 FP32 partial products
    // (launches with few output tiles split K over MX_SPLIT_MAX or fewer
    // work-groups per tile and add the partials in a fixed order); null:
    // never split.  splitk is set by the launcher.
    float* part;
    long partcap;  // floats in part
    int splitk;
};
constexpr int MX_SPLIT_MAX = 4;
// prec: MxPrec; dir 0: along x (i), 1: along y (j); epi: MxEpi
void launch_mx_gemm(int prec, int dir, int epi, const MxGemmArgs& a, int nbatch, hipStream_t s);

}  // namespace ibpm
