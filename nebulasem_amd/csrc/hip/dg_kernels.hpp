// Launchers of the DG Euler kernels (definitions in the .hip files).
#pragma once
#include <hip/hip_runtime.h>

#include "dg_common.hpp"

namespace nsem {

struct SplitState {
    const double* rho;
    const double* ux;
    const double* uy;
    const double* uz;
    const double* T;
};
struct SplitOut {
    double* ux;
    double* uy;
    double* uz;
    double* T;
    double* p;
};

// One fused residual + SSP-RK3 stage.  q: FP64 base state (5 SoA arrays:
// rho', m_x, m_y, m_z, Theta'); din: stage increment in (null at stage 1);
// out: b*(din + dt L(q + din)) to dout, or q + that to qout when last.
// The tangent-linear instantiation (residual type Dual<1>, S = double) also
// reads tq (tangent of q), tdin and writes tdout or tqout.
template <typename S>
struct RkArgs {
    const double* q[5];
    const S* din[5];
    S* dout[5];
    double* qout[5];
    double b;
    double bdt;   // b * dt (the factor of L in the output)
    int epb;      // elements per work-group (set by the launcher)
    int last;
    int check;
    int* flag;
    const double* tq[5];
    const double* tdin[5];
    double* tdout[5];
    double* tqout[5];
    // artificial viscosity: the stage state's strong gradients, [12][nb] in
    // the residual type (3a+b = d_a U_b, 9+a = d_a theta'), from
    // launch_mol_gradient; null without diffusion
    const void* grad;
};

// Strong gradients of U and theta' at the stage state q + din of the
// method-of-lines form (the split step's gradf<true> rule, evaluated in the
// residual type R), written to grad [12][nb] of R.
template <typename R, typename S>
hipError_t launch_mol_gradient(const DgDev& d, const double* const q[5], const S* const din[5], R* grad,
                               hipStream_t st);
// Vector-Jacobian product of the method-of-lines residual at Q = q + din:
// t = (dL/dQ)^T w, all FP64 SoA node arrays (din may be null).
// Epilogue (the adjoint's stage algebra, fused): with t = J^T w at a node,
//   t_out = t                                  (when t_out is set)
//   ad_out = c1 (w + dt t)                     (when ad_out is set)
//   comb_out = k0 y0 + k1 y1 + w + dt t        (when comb_out is set)
// ad_out and comb_out must not alias w (other work-groups read w); comb_out
// may alias y0 or y1 (each node reads its own entry before writing it).
struct VjpArgs {
    int epb;      // elements per work-group (set by the launcher)
    const double* q[5];
    const double* din[5];
    const double* w[5];
    double* t[5];
    double* ad_out[5];
    const double* y0[5];
    const double* y1[5];
    double* comb_out[5];
    double dt, c1, k0, k1;
};
hipError_t launch_rk_vjp(const DgDev& d, const VjpArgs& a, hipStream_t st);
size_t rk_vjp_lds_bytes(int nx, int ny, int nz, int epb, bool prims_in_lds);


// LDS of the RK stage kernel, in values of the residual type: per element
// max(15 flux rows x NP, 10 padded primitive rows + 5 face-flux rows x NFN),
// plus the three derivative matrices once per work-group.
__host__ __device__ inline int rk_prim_stride(int np) { return np + (np + 7) / 8; }
__host__ __device__ inline int rk_prim_slot(int node) { return node + (node >> 3); }
__host__ __device__ inline int rk_lds_per_element(int nx, int ny, int nz) {
    const int np = nx * ny * nz, nfn = 2 * (nx * ny + nx * nz + ny * nz);
    const int a = 15 * np, b = 10 * rk_prim_stride(np) + 5 * nfn;
    return a > b ? a : b;
}
inline size_t rk_lds_bytes(int nx, int ny, int nz, int epb, size_t value_size) {
    return ((size_t)epb * rk_lds_per_element(nx, ny, nz) + nx * nx + ny * ny + nz * nz) * value_size;
}

template <typename R, typename S>
hipError_t launch_rk_stage(const DgDev& d, const RkArgs<S>& a, hipStream_t st);

// Diagnostics input: the split (primitive) state or the method-of-lines
// conserved state, whichever is current.
struct DiagState {
    int mol;
    const double* rho;
    const double* ux;
    const double* uy;
    const double* uz;
    const double* T;
    const double* p;
    const double* q[5];
};
constexpr int kDiagBlocks = 1024;
// out (device, 6 doubles): Courant max, min, avg, mass, energy, volume;
// slots: kDiagBlocks * 6 doubles.
hipError_t launch_diagnostics(const DgDev& d, const DiagState& s, const double* gh, double* slots,
                              double* out, hipStream_t st);

// One split forward-Euler step (two fused element kernels; with
// d.diffusion a gradient kernel first, writing grad, which d.grad reads).
// A NaN or Inf in the new state sets *d.flag.
hipError_t launch_split_step(const DgDev& d, const SplitState& s, double* rho1, const SplitOut& o,
                             double* grad, hipStream_t st);

// Conserved perturbations (rho', m_x, m_y, m_z, Theta'), SoA node arrays.
struct ConsOut {
    double* q[5];
};
struct ConsIn {
    const double* q[5];
};
// Primitive split state -> conserved perturbations (the form set_state uses).
hipError_t launch_prim_to_cons(const DgDev& d, const SplitState& s, const ConsOut& q, hipStream_t st);
// Conserved perturbations -> rho, U, theta', p' (p' without cancellation).
hipError_t launch_cons_to_prim(const DgDev& d, const ConsIn& q, double* rho, const SplitOut& o, hipStream_t st);

}  // namespace nsem
