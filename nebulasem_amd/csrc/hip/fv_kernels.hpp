// FV Poisson operator and CG kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nsem {

// Symmetric FV operator A = diag - offdiag on interior cells.
template <typename R>
struct FvOp {
    uint64_t n = 0;
    int structured = 0;        // block-grid numbering: only +x/+y/+z couplings stored
    int nx = 0, ny = 0, nz = 0;
    int64_t sx = 0, sy = 0, sz = 0;
    const R* diag = nullptr;   // [n]
    const R* idiag = nullptr;  // [n] 1/diag
    const R* cpl = nullptr;    // [3][n] structured couplings (0 at walls)
    const uint32_t* nbr = nullptr;  // [6][n] neighbour or 0xffffffff (unstructured)
    const R* off = nullptr;         // [6][n]
    // structured with every x (y, z) coupling bitwise equal (a uniform grid
    // and coefficient): the marching SpMV uses cu[d] instead of cpl[d][.]
    int uniform = 0;
    R cu[3] = {0, 0, 0};
};

// Device-resident CG scalars.
struct CgScalars {
    double alpha, beta, rz, pAp, bb, bnorm;
    double res_ref, res0_ref, res_true, rnorm0, tol;
    int32_t iters, max_it, done, breakdown, mode;  // mode 0: reference criterion, 1: |r|/|b|
    int32_t keep_going;  // 1: a non-positive p.Ap does not stop the iteration (the reference's loop)
};

struct FvDivInput {
    uint64_t ncells;
    const uint32_t *fi0, *fi1, *allFaces, *FO, *FN;
    const double* fN;   // per face xyz
    const double* fI;   // per face
    const double* U;    // interior cells xyz
    const int32_t* patch_of_ghost;
    const int32_t* bc_kind;    // per patch: 0 NEUMANN 1 SYMMETRY 2 DIRICHLET
    const double* bc_value;    // per patch xyz
    const double* su_bc;       // boundary-condition part of the source
};

constexpr int kCgSlots = 2048;  // per-block partial slots (x4 values)

template <typename R>
hipError_t fv_cg_init(const FvOp<R>& A, const R* b, const R* x, R* r, R* z, R* p, CgScalars* sc,
                      double* slots, hipStream_t s);
template <typename R>
hipError_t fv_cg_iterations(const FvOp<R>& A, int count, R* x, R* r, R* z, R* p, R* Ap, CgScalars* sc,
                            double* slots, hipStream_t s);
template <typename R>
hipError_t fv_norm2(uint64_t n, const R* v, double* out, double* slots, hipStream_t s);
hipError_t fv_residual64(const FvOp<double>& A, const double* b, const double* x, double* y, float* y32,
                         double* out, double* slots, hipStream_t s);
hipError_t fv_correct(uint64_t n, double* x, const float* e, hipStream_t s);
hipError_t fv_divu(const FvDivInput& in, double* su, hipStream_t s);
// U -= gradf(p) on the interior cells; gkind/goff per ghost cell (nghost)
hipError_t fv_gradcorr(const FvDivInput& in, const double* p, const int32_t* gkind, const double* goff,
                       const double* cV, uint64_t nghost, double* U, hipStream_t s);
template <typename R>
hipError_t fv_convert(uint64_t n, const double* a, R* b, hipStream_t s);
template <typename R>
hipError_t fv_apply(const FvOp<R>& A, const R* x, R* y, hipStream_t s);

}  // namespace nsem
