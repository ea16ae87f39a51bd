// Conversions between the two device states of a DG Euler problem: the
// split step's primitive state (rho, U, theta', p', as apps/euler/euler.cpp
// stores it) and the method-of-lines conserved perturbations over the
// hydrostatic background (rho' = rho - rho_ref, m = rho U,
// Theta' = rho theta - rho_ref T0).  Each entry point that advances one form
// converts from the other on the device when the other was advanced last.
#include "dg_kernels.hpp"

namespace nsem {

__global__ void __launch_bounds__(256) prim_to_cons_kernel(DgDev d, SplitState s, ConsOut q) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.nb) return;
    const double rho = s.rho[n], rr = d.rho_ref[n];
    q.q[0][n] = rho - rr;
    q.q[1][n] = rho * s.ux[n];
    q.q[2][n] = rho * s.uy[n];
    q.q[3][n] = rho * s.uz[n];
    q.q[4][n] = rho * (s.T[n] + d.T0) - rr * d.T0;
}

// p' = p_ref ((Theta/Theta_ref)^gamma - 1) = p_ref expm1(gamma log1p(Theta'/Theta_ref)),
// the difference p - p_ref without cancellation.
__global__ void __launch_bounds__(256) cons_to_prim_kernel(DgDev d, ConsIn q, SplitOut o, double* rho_out) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.nb) return;
    const double rr = d.rho_ref[n];
    const double rho = rr + q.q[0][n];
    rho_out[n] = rho;
    o.ux[n] = q.q[1][n] / rho;
    o.uy[n] = q.q[2][n] / rho;
    o.uz[n] = q.q[3][n] / rho;
    o.T[n] = (q.q[4][n] + rr * d.T0) / rho - d.T0;
    o.p[n] = d.p_ref[n] * expm1(d.gamma * log1p(q.q[4][n] / (rr * d.T0)));
}

hipError_t launch_prim_to_cons(const DgDev& d, const SplitState& s, const ConsOut& q, hipStream_t st) {
    const unsigned blocks = (unsigned)((d.nb + 255) / 256);
    hipLaunchKernelGGL(prim_to_cons_kernel, dim3(blocks), dim3(256), 0, st, d, s, q);
    return hipGetLastError();
}

hipError_t launch_cons_to_prim(const DgDev& d, const ConsIn& q, double* rho, const SplitOut& o, hipStream_t st) {
    const unsigned blocks = (unsigned)((d.nb + 255) / 256);
    hipLaunchKernelGGL(cons_to_prim_kernel, dim3(blocks), dim3(256), 0, st, d, q, o, rho);
    return hipGetLastError();
}

}  // namespace nsem
