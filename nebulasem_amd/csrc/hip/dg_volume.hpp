// The weak-form DG volume term in the reference's accumulation order,
// shared by the split step and the field API's divergence kernel.
#pragma once
#include "dg_common.hpp"

namespace nsem {

// Jin's components in the LDS row order volume_reference_order reads: rows
// 3d .. 3d+2 hold column d of Jin (x: XX YX ZX, y: XY YY ZY, z: XZ YZ ZZ),
// components numbered XX0 YY1 ZZ2 XY3 YZ4 XZ5 YX6 ZY7 ZX8.
constexpr int kJinRow[9] = {0, 6, 8, 3, 1, 7, 5, 4, 2};

// Volume term of the reference's weak divergence (field.h:3443-3464) at
// target node (i,j,k), subtracted from r in the reference's accumulation
// order.  The reference walks source nodes s lexicographically and, at each,
// subtracts dot(F_s, Jin_s grad psi_t(x_s)) from every target on the three
// GLL lines through s.  Seen from the target the sources come as: x-line
// a < i, y-line b < j, z-line c < k, itself, z-line c > k, y-line b > j,
// x-line a > i.  Off the diagonal grad psi has one non-zero component, so
// Jin grad psi is one column of Jin times a derivative value, exactly as the
// reference's 3x3 product evaluates it.  F: LDS rows x 3 x NP; J: 9 x NP in
// the kJinRow order; D0, D1, D2: the 1-D derivative matrices.
// The walk itself: sub(s, w0, w1, w2) is called for each source node s of
// target (i,j,k) in the reference's order with w = Jin_s grad psi_t(x_s).
template <typename Sub>
__device__ __forceinline__ void volume_reference_walk(Sub&& sub, const double* J, int NP, int i, int j, int k,
                                                      int NX, int NY, int NZ, const double* D0, const double* D1,
                                                      const double* D2) {
    const int syz = NY * NZ;
    const int t = i * syz + j * NZ + k;
    for (int a = 0; a < i; a++) {
        const int s = a * syz + j * NZ + k;
        const double D = D0[a * NX + i];
        sub(s, J[s] * D, J[NP + s] * D, J[2 * NP + s] * D);
    }
    {
        const double d0 = D0[i * NX + i], d1 = D1[j * NY + j], d2 = D2[k * NZ + k];
        const double s0 = J[0 * NP + t] * d0 + J[3 * NP + t] * d1 + J[6 * NP + t] * d2;
        const double s1 = J[1 * NP + t] * d0 + J[4 * NP + t] * d1 + J[7 * NP + t] * d2;
        const double s2 = J[2 * NP + t] * d0 + J[5 * NP + t] * d1 + J[8 * NP + t] * d2;
        // y-line b < j, z-line c < k, the node itself, z-line c > k, y-line
        // b > j: NY + NZ - 1 sources for every node, walked by one loop whose
        // trip count does not depend on (j, k), so the lanes of a wavefront
        // stay converged; each step selects its source, column and
        // derivative value (at the node itself: c = k, the products above)
        const int base = i * syz;
#pragma unroll 1
        for (int m = 0; m < NY + NZ - 1; m++) {
            const bool ylo = m < j;
            const bool isy = ylo || m >= j + NZ;
            const int b = ylo ? m : m - NZ + 1;
            const int src = isy ? base + b * NZ + k : base + j * NZ + (m - j);
            const double D = isy ? D1[b * NY + j] : D2[(m - j) * NZ + k];
            const double* Jc = J + (isy ? 3 * NP : 6 * NP) + src;
            double w0 = Jc[0] * D, w1 = Jc[NP] * D, w2 = Jc[2 * NP] * D;
            if (m == j + k) { w0 = s0; w1 = s1; w2 = s2; }
            sub(src, w0, w1, w2);
        }
    }
    for (int a = i + 1; a < NX; a++) {
        const int s = a * syz + j * NZ + k;
        const double D = D0[a * NX + i];
        sub(s, J[s] * D, J[NP + s] * D, J[2 * NP + s] * D);
    }
}

template <int ROWS, int VECMASK>
__device__ __forceinline__ void volume_reference_order(double* r, const double* F, const double* J, int NP,
                                                       int i, int j, int k, int NX, int NY, int NZ,
                                                       const double* D0, const double* D1,
                                                       const double* D2) {
    volume_reference_walk(
        [&](int s, double w0, double w1, double w2) {
            for (int row = 0; row < ROWS; row++) {
                const double f0 = F[(row * 3 + 0) * NP + s];
                const double f1 = F[(row * 3 + 1) * NP + s];
                const double f2 = F[(row * 3 + 2) * NP + s];
                // vector flux: right-nested dot; tensor flux: row dot left to right
                const double c =
                    ((VECMASK >> row) & 1) ? dot3r(f0, f1, f2, w0, w1, w2) : f0 * w0 + f1 * w1 + f2 * w2;
                r[row] -= c;
            }
        },
        J, NP, i, j, k, NX, NY, NZ, D0, D1, D2);
}

}  // namespace nsem
