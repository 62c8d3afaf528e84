// adjoint.h -- kernels of the tangent-linear and adjoint advance! steps
// (adjoint.hip): the transposes of the multigrid transfer operators, the
// linearised nonlinear term and its transposes, and small reductions.
// Every launch works on one grid level unless stated; fields use the
// reference's layouts (an (nx-1) x (ny-1) circulation array, a flux array of
// (nx+1) x ny x-fluxes then nx x (ny+1) y-fluxes, a boundary vector of
// 2(nx+1) + 2(ny+1) entries).
#pragma once
#include "kernels.h"

namespace ibpm {

// Q = avg_flux(q + q0) (nonlinear.jl:17-38); q0 may be null
void launch_adj_avg_flux(const double* q, const double* q0, double* Q, const GridDev& g, hipStream_t s);
// q̄ (+)= avgᵀ Q̄ (acc: add into q̄, else overwrite)
void launch_adj_avg_flux_T(const double* Qb, double* qb, bool acc, const GridDev& g, hipStream_t s);

// fq (+)= direct_product(Q, Γ with its boundary ring 0.25 get_bc!(Γc)), the
// product of nonlinear.jl:73-114 (Γc null: zero ring, the coarsest level)
void launch_adj_dprod(const double* Q, const double* G, const double* Gc, double* fq, bool acc, const GridDev& g,
                      hipStream_t s);
// transposes of the product: Q̄ = ∂(fq)/∂Q ᵀ f̄ (for fixed Γ and ring) and,
// for fixed Q, Γ̄ += interior part, rinḡ = boundary-ring part (nbc entries,
// indexed like Γbc; the 0.25 get_bc! factor is not applied)
void launch_adj_dprod_TQ(const double* fb, const double* G, const double* Gc, double* Qb, const GridDev& g,
                         hipStream_t s);
void launch_adj_dprod_TG(const double* fb, const double* Q, double* Gb, double* ringb, const GridDev& g,
                         hipStream_t s);

// out (+)= scale rot!(q) over the whole level (curl.jl:22-34)
void launch_adj_rot(const double* q, double* G, double scale, bool acc, const GridDev& g, hipStream_t s);
// q = scale curl!(ψ) of one level with zero boundary values (curl.jl:51-71), the transpose of rot!
void launch_adj_curl_sg(const double* psi, double* q, double scale, const GridDev& g, hipStream_t s);

// boundary-value transposes: bc̄ = (the ψbc part of the multigrid curl!,
// curl.jl:81-116)ᵀ q̄; bc̄ = fac (apply_bc!)ᵀ r̄ (mg-utils.jl:43-60), both
// overwrite bc̄; r̄c += fac (get_bc!)ᵀ bc̄ on the coarser level (mg-utils.jl:7-32)
void launch_adj_curl_bc_T(const double* qb, double* bcb, const GridDev& g, hipStream_t s);
void launch_adj_apply_bc_T(const double* rb, double fac, double* bcb, const GridDev& g, hipStream_t s);
void launch_adj_get_bc_T(const double* bcb, double fac, double* rcb, const GridDev& g, hipStream_t s);

// coarsify! (mg-utils.jl:71-106) transposed: Ḡ_fine += Rᵀ Ḡ_coarse(central
// block), then the central block of Ḡ_coarse is zeroed (two launches)
void launch_adj_coarsify_T(double* Gfb, double* Gcb, const CoarsifyArgs& r, hipStream_t s);

// y (+)= a x
void launch_adj_axpy(double* y, const double* x, double a, long n, hipStream_t s);
// x[0..n/2) += a, x[n/2..n) += b (the cotangent of the force sums)
void launch_adj_add_halves(double* x, int n, double a, double b, hipStream_t s);
// y = M x for a row-major n x n matrix M (row r at M + r n)
void launch_adj_gemv(const double* M, const double* x, double* y, int n, hipStream_t s);

// Reductions of q̄0 over every level of the MovingGrid base flux
// (timestepping.jl:284-317): out[0] = Σ hc q̄0x, out[1] = Σ hc q̄0y,
// out[2] = Σ hc (-YY q̄0x + XX q̄0y); and of δF̃b: out[3] = Σ δF̃b_x, out[4] = Σ δF̃b_y.
// The partial sums of each work-group go to part (3 per group), then a
// single work-group adds them in a fixed order.
void launch_adj_q0_reduce(const double* q0b, int mg, double offx, double offy, double* part, double* out,
                          const GridDev& g, hipStream_t s);
void launch_adj_sum_halves(const double* x, int n, double* out2, hipStream_t s);
int adj_q0_reduce_parts(const GridDev& g, int mg);

}  // namespace ibpm
