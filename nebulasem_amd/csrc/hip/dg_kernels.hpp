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

// One split forward-Euler step (two fused element kernels).
hipError_t launch_split_step(const DgDev& d, const SplitState& s, double* rho1, const SplitOut& o,
                             hipStream_t st);

}  // namespace nsem
