// Device-side data of a DG Euler problem and the helpers its kernels share.
//
// Layout in HBM (one GPU, interior nodes only; ghost values are never stored,
// they are produced from the owner state by the boundary condition when a
// face flux needs them):
//   node n = element*NP + i*NPY*NPZ + j*NPZ + k       (as the reference)
//   state      SoA double arrays  rho, ux, uy, uz, T (theta'), p (p')
//   metric     SoA double JcV[9][nb] = Jinv*cV, cV[nb]; per element
//              Jinv_e (9) and the cell volume when the elements are affine
//   background SoA double p_ref[nb], rho_ref[nb]
//   faces      per element, per face group g (0..5): EFace {facet, other, meta}
//              per facet: area normal gFN (3 doubles)
//   basis      per direction: LGL weights w[n], derivative D[s*n+i] = l'_i(x_s)
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nsem {

constexpr int kMaxN = 16;  // LGL points per direction supported

// meta bits: [0:3) other face group, [3] self is owner, [4] face present,
// [5] other is a ghost (boundary), [8:24) boundary patch index
// meta bits: 0-2 the other side's face group, 3 this element owns the face,
// 4 present, 5 boundary (ghost), 6 mortar side; bits 8+: the boundary
// patch, or for a mortar side its slot in DgDev::msurf.
struct EFace {
    uint32_t facet;
    uint32_t other;
    uint32_t meta;
};

// A 2:1 non-conforming face: a refined element's face lying on a split side
// of a coarse element (field.h:2018-2248).
struct MortarFace {
    uint32_t coarse;    // the coarse element
    int32_t side;       // its side (face group) holding the face
    int32_t h0, h1;     // the half of the side the face covers, along the side's two node directions
    int32_t fine_owner; // 1: the refined element owns the face (FMC 1), 0: the coarse one (FMC 2)
    int32_t n0, n1;     // node extents of the side
    int32_t pad;
    double nf[3];       // the face's area normal, pointing out of the owner
};

enum BcKind : int32_t { BC_NEUMANN = 0, BC_SYMMETRY = 1, BC_DIRICHLET = 2 };

constexpr int kMaxPatches = 64;

struct BcTable {
    int32_t kind[4][kMaxPatches];     // rho, U, T, p
    double value[6][kMaxPatches];     // rho, Ux, Uy, Uz, T, p
};

struct DgDev {
    int npx, npy, npz, np;
    uint32_t nelem;                   // interior elements
    uint64_t nb;                      // interior nodes
    double dt, P0, T0, R, gamma, cp;
    // binomial(gamma, k), k = 1..8: the FP64 series of (1 + x)^gamma - 1
    // (kernel arguments, so the kernels read them from scalar registers)
    double ppc[8];
    double g[3];
    int buoyancy;
    const double* JcV;                // [9][nb]
    const double* cV;                 // [nb]
    const float* JcV32;               // [9][nb] FP32 copies for the FP32 residual
    const float* cV32;                // [nb]
    const double* p_ref;              // [nb]
    const double* rho_ref;            // [nb]
    const EFace* eface;               // [nelem*6]
    const double* gFN;                // [nfacets*3]
    const double* D[3];               // [n*n] per direction
    const double* w[3];               // [n] per direction
    const BcTable* bc;
    // artificial viscosity of the split step (euler.cpp:188-190): mu =
    // rho visc, theta diffusivity mu iPr; grad [12][nb] holds the strong
    // gradients of U (3a+b = d_a U_b) and theta (9+a = d_a theta)
    int diffusion;
    double visc, iPr;
    const double* grad;
    // affine elements (every element's Jinv constant to 1e-10 relative): the
    // method-of-lines kernels form JcV = Jinv_e cV and cV = vol_e W(node)
    // from emet [nelem][10] (Jinv_e in tensor order, then the cell volume)
    // and wnode [NP] = w_i w_j w_k / 8 instead of reading the node arrays
    int affine;
    const double* emet;
    const double* wnode;
    // 2:1 non-conforming faces (nmort of them, sides of up to npfm nodes):
    // per face node, the refined side's node (mort_fine), the entry of that
    // node in msurf (mort_fms), the face-node weight w_a w_b / 4 (mort_w) and
    // the background density interpolated from the coarse side minus the
    // refined side's (mort_bg); psiRef / psiCor [2][n*n] per direction.  A
    // pre-pass of every stage writes the face flux per unit weight (mphi)
    // and the surface terms of every mortar side (msurf [slot][npfm][5]);
    // cslot [ncslot][8] = element, side, children, slot, child faces (4).
    int nmort, npfm, ncslot;
    const MortarFace* mort;
    const uint64_t* mort_fine;
    const uint32_t* mort_fms;
    const double* mort_w;
    const double* mort_bg;
    const double* psiRef[3];
    const double* psiCor[3];
    const uint32_t* cslot;
    double* mphi;
    double* msurf;
    // tangents of mphi / msurf (tangent-linear stages) and the mortar faces'
    // share of a vector-Jacobian product [5][nb] (adjoint)
    double* mphi_t;
    double* msurf_t;
    double* madj;
    // the split step's strong-gradient lift of the mortar sides [slot][npfm][12]
    double* mgrad;
    // boundary (ghost) sides of the vector-Jacobian product: ghost_rec
    // [nghost] = element * 6 + face group, ghost_of [nelem * 6] = the
    // record of a ghost side (or ~0u), gadj [nghost][5][npfm] the face-node
    // terms a pre-pass writes for the element kernel to add in face order
    int nghost;
    const uint32_t* ghost_rec;
    const uint32_t* ghost_of;
    double* gadj;
    // set to 1 by a kernel that produced a NaN or Inf in the state (the
    // split step's update; the RK stage kernels take it in RkArgs)
    int* flag;
};

// Node-array access.  With O32 the byte offset is formed in 32 bits (node
// arrays below 4 GiB each), so the load takes the array base from scalar
// registers and one shared 32-bit offset instead of a 64-bit address per
// array.
// (The empty asm keeps the 32-bit offset opaque, so it is not re-widened to
// a 64-bit address computation.)
__device__ __forceinline__ uint32_t byte_off(uint64_t i, uint32_t size) {
    uint32_t off = (uint32_t)i * size;
    __asm__("" : "+v"(off));
    return off;
}
template <bool O32, typename T>
__device__ __forceinline__ T ldn(const T* base, uint64_t i) {
    if constexpr (O32)
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off(i, sizeof(T)));
    else
        return base[i];
}
template <bool O32, typename T>
__device__ __forceinline__ void stv(T* base, uint64_t i, T v) {
    if constexpr (O32)
        *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off(i, sizeof(T))) = v;
    else
        base[i] = v;
}
__device__ __forceinline__ double dot3r(double a0, double a1, double a2, double b0, double b1,
                                        double b2) {
    // right-nested like the reference's unrolled vector dot product
    double r = a2 * b2 + 0.0;
    r = a1 * b1 + r;
    return a0 * b0 + r;
}

// SYMMETRY ghost velocity: tangential projection rescaled to |U|
// (src/tensor/tensor.h:483-492).
__device__ __forceinline__ void sym_vector(const double* p, const double* n, double* r) {
    double mn = sqrt(dot3r(n[0], n[1], n[2], n[0], n[1], n[2]));
    double e0 = n[0] / mn, e1 = n[1] / mn, e2 = n[2] / mn;
    double A0 = 1 - e0 * e0, A1 = 1 - e1 * e1, A2 = 1 - e2 * e2;
    double A3 = 0 - e0 * e1, A4 = 0 - e1 * e2, A5 = 0 - e0 * e2;
    r[0] = A0 * p[0] + A3 * p[1] + A5 * p[2];
    r[1] = A3 * p[0] + A1 * p[1] + A4 * p[2];
    r[2] = A5 * p[0] + A4 * p[1] + A2 * p[2];
    double mr = sqrt(dot3r(r[0], r[1], r[2], r[0], r[1], r[2]));
    if (fabs(mr) <= 1e-7) return;
    double s = sqrt(dot3r(p[0], p[1], p[2], p[0], p[1], p[2])) / mr;
    r[0] *= s;
    r[1] *= s;
    r[2] *= s;
}

// Ghost velocity of a boundary face node (applyExplicitBCs, field.h:2660-2690):
// SYMMETRY reflects to the tangential plane, DIRICHLET is the patch value,
// NEUMANN copies (|dv| is zero for DG ghost nodes).  Scalars are always
// zero-gradient copies on this path.
__device__ __forceinline__ void ghost_velocity(const BcTable& bt, int pi, const double* own,
                                               const double* n, double* out) {
    const int kind = bt.kind[1][pi];
    if (kind == BC_SYMMETRY) {
        sym_vector(own, n, out);
    } else if (kind == BC_DIRICHLET) {
        out[0] = bt.value[1][pi]; out[1] = bt.value[2][pi]; out[2] = bt.value[3][pi];
    } else {
        out[0] = own[0] + bt.value[1][pi] * 0.0;
        out[1] = own[1] + bt.value[2][pi] * 0.0;
        out[2] = own[2] + bt.value[3][pi] * 0.0;
    }
}

// Face membership of node (i,j,k) for face group g (dg.cpp face_map).
__device__ __forceinline__ bool on_face(int g, int i, int j, int k, int NX, int NY, int NZ) {
    switch (g) {
        case 0: return k == 0;
        case 1: return k == NZ - 1;
        case 2: return j == 0;
        case 3: return j == NY - 1;
        case 4: return i == 0;
        default: return i == NX - 1;
    }
}

// Face-local coordinates (a,b) of node (i,j,k) on face group g.
__device__ __forceinline__ void face_ab(int g, int i, int j, int k, int& a, int& b) {
    if (g < 2) { a = i; b = j; }
    else if (g < 4) { a = i; b = k; }
    else { a = j; b = k; }
}

// Node of element `cell` on face group g at face-local (a,b).
__device__ __forceinline__ uint64_t face_node(uint32_t cell, int g, int a, int b, int NX, int NY,
                                              int NZ) {
    const int v = (g & 1) ? (g < 2 ? NZ - 1 : g < 4 ? NY - 1 : NX - 1) : 0;
    int i, j, k;
    if (g < 2) { i = a; j = b; k = v; }
    else if (g < 4) { i = a; j = v; k = b; }
    else { i = v; j = a; k = b; }
    return (uint64_t)cell * (NX * NY * NZ) + (uint64_t)(i * NY * NZ + j * NZ + k);
}

// Facet-node weight w_a w_b / 4 in the owner's face layout (dg.cpp:368-404).
__device__ __forceinline__ double face_weight(const DgDev& d, int owner_g, int a, int b) {
    if (owner_g < 2) return d.w[0][a] * d.w[1][b] / 4;
    if (owner_g < 4) return d.w[0][a] * d.w[2][b] / 4;
    return d.w[1][a] * d.w[2][b] / 4;
}

// Face-node task t (0 .. NFN-1) of an element: face group g, face-local
// (a, b) and the element-local node.  Groups 0,1 (k = 0, NZ-1) have NX*NY
// nodes, 2,3 (j) NX*NZ, 4,5 (i) NY*NZ, enumerated group by group in the
// (a, b) order of face_ab.
__device__ __forceinline__ void face_task(int t, int NX, int NY, int NZ, int& g, int& fa, int& fb,
                                          int& node) {
    const int FK = NX * NY, FJ = NX * NZ, FI = NY * NZ;
    int i, j, k, l;
    if (t < 2 * FK) {
        g = t >= FK; l = t - g * FK;
        fa = l / NY; fb = l - fa * NY;
        i = fa; j = fb; k = g ? NZ - 1 : 0;
    } else if (t < 2 * (FK + FJ)) {
        l = t - 2 * FK;
        const int h = l >= FJ;
        g = 2 + h; l -= h * FJ;
        fa = l / NZ; fb = l - fa * NZ;
        i = fa; j = h ? NY - 1 : 0; k = fb;
    } else {
        l = t - 2 * (FK + FJ);
        const int h = l >= FI;
        g = 4 + h; l -= h * FI;
        fa = l / NZ; fb = l - fa * NZ;
        i = h ? NX - 1 : 0; j = fa; k = fb;
    }
    node = i * NY * NZ + j * NZ + k;
}

// Inverse of face_task for a node on face group g.
__device__ __forceinline__ int face_task_index(int g, int fa, int fb, int NX, int NY, int NZ) {
    const int FK = NX * NY, FJ = NX * NZ, FI = NY * NZ;
    switch (g) {
        case 0: return fa * NY + fb;
        case 1: return FK + fa * NY + fb;
        case 2: return 2 * FK + fa * NZ + fb;
        case 3: return 2 * FK + FJ + fa * NZ + fb;
        case 4: return 2 * (FK + FJ) + fa * NZ + fb;
        default: return 2 * (FK + FJ) + FI + fa * NZ + fb;
    }
}

}  // namespace nsem
