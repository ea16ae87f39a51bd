// The device-side state of a DG Euler problem (the nsem_dg handle), shared
// by the files that implement its entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../status.hpp"
#include "dg_common.hpp"
#include "dg_kernels.hpp"
#include "nebulasem_amd.h"

#define HIP_TRY(x)                                                                  \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess)                                                       \
            return nsem_fail(NSEM_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

namespace nsem {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { bytes = n; return hipMalloc(&p, n ? n : 8); }
    template <typename T> T* as() const { return (T*)p; }
};

}  // namespace nsem

struct nsem_dg {
    using DevBuf = nsem::DevBuf;
    using DgDev = nsem::DgDev;
    DgDev dev{};
    nsem_euler_params prm{};
    int device = 0;  // the HIP device current at nsem_dg_create
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t nb = 0;
    // geometry
    DevBuf JcV, cV, JcV32, cV32, p_ref, rho_ref, eface, gFN, D[3], w[3], bc, emet, wnode;
    // 2:1 non-conforming faces (empty on conforming meshes)
    DevBuf mort, mort_fine, mort_fms, mort_w, mort_bg, psiRef[3], psiCor[3], cslot, mphi, msurf, mphi_t, msurf_t,
        madj, mgrad;
    // ghost sides of the vector-Jacobian product
    DevBuf ghost_rec, ghost_of, gadj;
    // double-buffered state: rho ux uy uz T p
    DevBuf st[2][6];
    int cur = 0;
    // method-of-lines state: FP64 conserved perturbations (rho', m, Theta'),
    // double-buffered, and stage increments in FP64 and FP32
    DevBuf mq[2][5], d64[2][5], d32[2][5], flag;
    // diagnostics: g.x per node, block partials, result
    DevBuf gh, diag_slots, diag_out;
    // strong gradients of U and theta of the viscous split step [12][nb]
    DevBuf grad;
    // workspace of the derivative entry points (grown on demand, kept)
    DevBuf adj_ws;
    int mcur = 0;
    // which of the two forms holds the current state: both after set_state,
    // the one advanced last otherwise (the other is rebuilt from it on the
    // device before it is advanced or read)
    bool split_valid = true;
    bool mol_valid = true;
    // profiling of the dominant kernels
    bool profile = false;
    double prof_ms[4] = {0, 0, 0, 0};
    uint64_t prof_n[4] = {0, 0, 0, 0};
    std::vector<double> host_p_ref, host_rho_ref;
    ~nsem_dg() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// nsteps FP64 SSP-RK3 steps of the method-of-lines state, queued on the
// handle's stream without waiting or checking the non-finite flag.
int nsem_rk3_fp64_queue(nsem_dg* d, int32_t nsteps);

// Make the method-of-lines (conserved) state current, converting from the
// split state on the device when the split step ran last; likewise the
// split state.  Both queue on the handle's stream.
int nsem_dg_require_mol(nsem_dg* d);
int nsem_dg_require_split(nsem_dg* d);
// Wait for the handle's stream, then read and clear the non-finite flag:
// NSEM_ENONFINITE (naming `what`) when a kernel raised it, else NSEM_OK.
int nsem_dg_take_flag(nsem_dg* d, const char* what);

// The current state (split or method-of-lines form) as the diagnostics
// kernels read it.
inline nsem::DiagState nsem_dg_diag_state(const nsem_dg* d) {
    nsem::DiagState ds{};
    ds.mol = d->split_valid ? 0 : 1;
    ds.rho = d->st[d->cur][0].as<double>();
    ds.ux = d->st[d->cur][1].as<double>();
    ds.uy = d->st[d->cur][2].as<double>();
    ds.uz = d->st[d->cur][3].as<double>();
    ds.T = d->st[d->cur][4].as<double>();
    ds.p = d->st[d->cur][5].as<double>();
    for (int v = 0; v < 5; v++) ds.q[v] = d->mq[d->mcur][v].as<double>();
    return ds;
}

// NSEM_OK, or NSEM_EUNSUPPORTED when the problem has artificial viscosity,
// which the derivative (tangent-linear, adjoint) entry points do not carry.
int nsem_mol_supported(const nsem_dg* d);
