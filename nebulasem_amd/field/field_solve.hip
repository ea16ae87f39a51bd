// SolveT (src/solvers/solve.cpp:39-559) for FV systems of any value type,
// every method (Jacobi, SOR, conjugate gradient for symmetric systems,
// biconjugate gradient otherwise) and preconditioner (none, DIAG, SSOR,
// DILU), in the reference's order of operations:
//  * the forward / backward substitutions and the SOR sweep are sequential
//    in cell order in the reference; here each runs as a sequence of
//    level-set launches (cells whose earlier neighbours are all done), and
//    every cell sums its faces in the reference's order, so the results are
//    those of the sequential loop;
//  * dot products and residual norms are summed in index order by one
//    work-group (products in parallel, then one lane per component);
//  * the scalars alpha, beta and the convergence test are formed on the
//    host from those sums, as the reference forms them.
// The components of a vector system are independent lanes of the same
// operations.  Every kernel is bounded by its launch size; the iteration
// count by max_iterations.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "field.h"
#include "field_mesh.h"

namespace nsem_field {
namespace {

constexpr int kT = 256;
constexpr int kMaxNc = 9;
inline unsigned nblk(size_t n) { return (unsigned)((n + kT - 1) / kT); }

struct Lanes {
    double v[kMaxNc];
};

struct Faces {
    const uint32_t *FO, *FN, *ent_off;
    const uint64_t* ent;
};
Faces faces() {
    const MeshTables& m = mesh_tables();
    return Faces{m.FO, m.FN, m.ent_off, m.ent};
}

// ------------------------------------------------------------- kernels --
// r = q ap over every node, then r -= q[other] c over the faces of each
// interior node's cell (mul: c0 = ann, c1 = ano; mult: c0 = ano, c1 = ann)
__global__ void spmv_kernel(Faces f, int nc, const double* q, const double* ap, const double* c0, const double* c1,
                            double* r, uint64_t nall, uint64_t nbf) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    for (int c = 0; c < nc; c++) {
        double acc = q[i * nc + c] * ap[i];
        if (i < nbf)
            for (uint32_t e = f.ent_off[i]; e < f.ent_off[i + 1]; e++) {
                const uint64_t v = f.ent[e], k = v >> 1;
                if (v & 1) acc -= q[(uint64_t)f.FO[k] * nc + c] * c1[k];
                else acc -= q[(uint64_t)f.FN[k] * nc + c] * c0[k];
            }
        r[i * nc + c] = acc;
    }
}

// r = Su over every node, then r += q[other] c (getRHS: c0 = ann, c1 = ano)
__global__ void rhs_kernel(Faces f, int nc, const double* q, const double* Su, const double* c0, const double* c1,
                           double* r, uint64_t nall, uint64_t nbf) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    for (int c = 0; c < nc; c++) {
        double acc = Su[i * nc + c];
        if (i < nbf)
            for (uint32_t e = f.ent_off[i]; e < f.ent_off[i + 1]; e++) {
                const uint64_t v = f.ent[e], k = v >> 1;
                if (v & 1) acc += q[(uint64_t)f.FO[k] * nc + c] * c1[k];
                else acc += q[(uint64_t)f.FN[k] * nc + c] * c0[k];
            }
        r[i * nc + c] = acc;
    }
}

// out[c] = sum over i < n of x[i,c] y[i,c], added in index order
__global__ void __launch_bounds__(kT) seq_dot_kernel(int nc, const double* x, const double* y, uint64_t n,
                                                     double* out) {
    __shared__ double prod[kT * kMaxNc];
    double acc = 0;
    for (uint64_t base = 0; base < n; base += kT) {
        const uint64_t i = base + threadIdx.x;
        if (i < n)
            for (int c = 0; c < nc; c++) prod[threadIdx.x * nc + c] = x[i * nc + c] * y[i * nc + c];
        __syncthreads();
        if ((int)threadIdx.x < nc) {
            const int c = threadIdx.x;
            const uint64_t m = n - base < (uint64_t)kT ? n - base : (uint64_t)kT;
            for (uint64_t j = 0; j < m; j++) acc += prod[j * nc + c];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < nc) out[threadIdx.x] = acc;
}

// Y = I + X alpha (componentwise alpha) over the interior nodes
__global__ void axpy_kernel(int nc, double* Y, const double* I, const double* X, Lanes alpha, uint64_t nbf) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbf) return;
    for (int c = 0; c < nc; c++) Y[i * nc + c] = I[i * nc + c] + X[i * nc + c] * alpha.v[c];
}

// Z = R iD over the interior nodes
__global__ void diag_kernel(int nc, double* Z, const double* R, const double* iD, uint64_t nbf) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbf) return;
    for (int c = 0; c < nc; c++) Z[i * nc + c] = R[i * nc + c] * iD[i];
}

// Z = Z D over every node
__global__ void scale_kernel(int nc, double* Z, const double* D, uint64_t nall) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    for (int c = 0; c < nc; c++) Z[i * nc + c] = Z[i * nc + c] * D[i];
}

// cF = iD r over every node (the Jacobi sweep)
__global__ void jacobi_kernel(int nc, double* cF, const double* iD, const double* r, uint64_t nall) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    for (int c = 0; c < nc; c++) cF[i * nc + c] = iD[i] * r[i * nc + c];
}

// AP = cF - p over the interior nodes
__global__ void diff_kernel(int nc, double* AP, const double* cF, const double* p, uint64_t nbf) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbf) return;
    for (int c = 0; c < nc; c++) AP[i * nc + c] = cF[i * nc + c] - p[i * nc + c];
}

__global__ void zero_kernel(double* x, uint64_t from, uint64_t to) {
    const uint64_t i = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < to) x[i] = 0;
}

// one level of the forward (fw) or backward substitution (Substitute_,
// solve.cpp:129-197) for the cells of the level: X[i] = (B[i] + sum over
// the cell's faces towards earlier cells of X[other] a) iD[i], a = ann
// (owner side) / ano (neighbour side), swapped for the transposed system
__global__ void subst_kernel(Faces f, int nc, const uint32_t* cells, uint32_t ncell, int fw, int tr, double* X,
                             const double* B, const double* iD, const double* ano, const double* ann) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const uint32_t i = cells[q];
    for (int c = 0; c < nc; c++) {
        double ncF = B[(uint64_t)i * nc + c];
        double val = 0;
        for (uint32_t e = f.ent_off[i]; e < f.ent_off[i + 1]; e++) {
            const uint64_t v = f.ent[e], k = v >> 1;
            const uint32_t c1 = f.FO[k], c2 = f.FN[k];
            if (!(v & 1)) {
                if ((fw && c2 < c1) || (!fw && c1 < c2)) val += X[(uint64_t)c2 * nc + c] * (tr ? ano[k] : ann[k]);
            } else {
                if ((fw && c2 > c1) || (!fw && c1 > c2)) val += X[(uint64_t)c1 * nc + c] * (tr ? ann[k] : ano[k]);
            }
        }
        ncF += val;
        ncF *= iD[i];
        X[(uint64_t)i * nc + c] = ncF;
    }
}

// one level of the SOR sweep (Sweep_, solve.cpp:64-103):
// X[i] = X[i] (1 - w) + (B[i] + sum X[other] a) iD[i] w
__global__ void sweep_kernel(Faces f, int nc, const uint32_t* cells, uint32_t ncell, double* X, const double* B,
                             const double* iD, const double* ano, const double* ann, double omega) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell) return;
    const uint32_t i = cells[q];
    for (int c = 0; c < nc; c++) {
        double ncF = B[(uint64_t)i * nc + c];
        for (uint32_t e = f.ent_off[i]; e < f.ent_off[i + 1]; e++) {
            const uint64_t v = f.ent[e], k = v >> 1;
            if (v & 1) ncF += X[(uint64_t)f.FO[k] * nc + c] * ano[k];
            else ncF += X[(uint64_t)f.FN[k] * nc + c] * ann[k];
        }
        ncF *= iD[i];
        X[(uint64_t)i * nc + c] = X[(uint64_t)i * nc + c] * (1 - omega) + ncF * (omega);
    }
}

// ---------------------------------------------------- level schedules --
// cells grouped by level: a cell's level is one more than the highest
// level among its neighbours that come before it in the sequential order
struct Schedule {
    std::vector<uint32_t> off;  // level l: cells[off[l] .. off[l+1])
    uint32_t* cells = nullptr;
    size_t n = 0;
};
struct Schedules {
    uint64_t generation = ~0ull;
    Schedule fw, bw, sor;
};
Schedules& schedules() {
    static auto* s = new Schedules();
    return *s;
}

void build_schedule(Schedule& s, const std::vector<uint32_t>& order) {
    const MeshTables& m = mesh_tables();
    const size_t n = order.size();
    std::vector<uint32_t> pos(n), lvl(n, 0);
    for (size_t q = 0; q < n; q++) pos[order[q]] = (uint32_t)q;
    uint32_t maxl = 0;
    for (size_t q = 0; q < n; q++) {
        const uint32_t i = order[q];
        uint32_t l = 0;
        for (uint32_t e = m.ent_off_h[i]; e < m.ent_off_h[i + 1]; e++) {
            const uint64_t v = m.ent_h[e], k = v >> 1;
            const uint32_t o = (v & 1) ? m.FO_h[k] : m.FN_h[k];
            if (o < n && pos[o] < q) l = std::max(l, lvl[o] + 1);
        }
        lvl[i] = l;
        maxl = std::max(maxl, l);
    }
    s.off.assign(maxl + 2, 0);
    for (size_t i = 0; i < n; i++) s.off[lvl[i] + 1]++;
    for (uint32_t l = 0; l <= maxl; l++) s.off[l + 1] += s.off[l];
    std::vector<uint32_t> cells(n), fill(s.off.begin(), s.off.end() - 1);
    // within a level, cells keep their sequential order
    for (size_t q = 0; q < n; q++) cells[fill[lvl[order[q]]]++] = order[q];
    if (s.cells) dev_free(s.cells, s.n * sizeof(uint32_t));
    s.n = n;
    s.cells = static_cast<uint32_t*>(dev_alloc(n * sizeof(uint32_t)));
    upload(s.cells, cells.data(), n * sizeof(uint32_t));
}

const Schedules& level_schedules() {
    const MeshTables& m = mesh_tables();
    Schedules& s = schedules();
    if (s.generation == m.generation) return s;
    const uint32_t n = m.nbcs;
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    build_schedule(s.fw, order);
    for (uint32_t i = 0; i < n; i++) order[i] = n - 1 - i;
    build_schedule(s.bw, order);
    // the SOR sweep visits the cells without boundary faces first, then
    // those with (ForwardSweep, solve.cpp:104-120)
    order.clear();
    for (uint32_t i = 0; i < n; i++)
        if (!m.has_boundary_h[i]) order.push_back(i);
    for (uint32_t i = 0; i < n; i++)
        if (m.has_boundary_h[i]) order.push_back(i);
    build_schedule(s.sor, order);
    s.generation = m.generation;
    return s;
}

// ------------------------------------------------------------- solver --
struct Buf {
    double* p = nullptr;
    size_t bytes = 0;
    explicit Buf(size_t n) : bytes(n * sizeof(double)) {
        p = static_cast<double*>(dev_alloc(bytes));
        check(hipMemsetAsync(p, 0, bytes, stream()), "solver buffer");
    }
    ~Buf() { dev_free(p, bytes); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
};

double mag_lanes(const Lanes& a, int nc) {
    if (nc == 1) return std::fabs(a.v[0]);
    double s = 0;
    for (int c = nc - 1; c >= 0; c--) s = a.v[c] * a.v[c] + s;
    return std::sqrt(s);
}
double sdiv(double p, double q) { return q ? (p / q) : 0; }

class SolverT {
   public:
    SolverT(const double* ap, const double* ano, const double* ann, const double* Su, double* cF, int nc, int flags)
        : m_(mesh_tables()), f_(faces()), ap_(ap), ano_(ano), ann_(ann), Su_(Su), cF_(cF), nc_(nc), flags_(flags),
          nall_(m_.nall), nbf_(m_.nbf), r_(nc * (nall_ + 1)), p_(nc * (nall_ + 1)), AP_(nc * (nall_ + 1)),
          D_(nall_ + 1), iD_(nall_ + 1), dot_(kMaxNc) {}

    void run();

   private:
    const MeshTables& m_;
    Faces f_;
    const double *ap_, *ano_, *ann_, *Su_;
    double* cF_;
    int nc_, flags_;
    uint64_t nall_, nbf_;
    Buf r_, p_, AP_, D_, iD_, dot_;
    std::unique_ptr<Buf> r1_, p1_, AP1_;

    bool symmetric() const { return flags_ & 1; }
    Lanes dot(const double* x, const double* y) {
        hipLaunchKernelGGL(seq_dot_kernel, dim3(1), dim3(kT), 0, stream(), nc_, x, y, nbf_, dot_.p);
        check(hipGetLastError(), "dot");
        Lanes out{};
        download(out.v, dot_.p, nc_ * sizeof(double));
        return out;
    }
    double residual(const double* r) {
        const Lanes a = dot(r, r), b = dot(cF_, cF_);
        return std::sqrt(sdiv(mag_lanes(a, nc_), mag_lanes(b, nc_)));
    }
    void axpy(double* Y, const double* I, const double* X, const Lanes& a) {
        hipLaunchKernelGGL(axpy_kernel, dim3(nblk(nbf_)), dim3(kT), 0, stream(), nc_, Y, I, X, a, nbf_);
    }
    void spmv(const double* q, double* r, bool transpose) {
        hipLaunchKernelGGL(spmv_kernel, dim3(nblk(nall_)), dim3(kT), 0, stream(), f_, nc_, q, ap_,
                           transpose ? ano_ : ann_, transpose ? ann_ : ano_, r, nall_, nbf_);
    }
    void copy(double* dst, const double* src) {
        copy_d2d(dst, src, (size_t)nc_ * nall_ * sizeof(double));
    }
    void zero_ghosts(double* x) {
        const uint64_t from = nbf_ * nc_, to = nall_ * nc_;
        if (to > from)
            hipLaunchKernelGGL(zero_kernel, dim3(nblk(to - from)), dim3(kT), 0, stream(), x, from, to);
    }
    void substitute(const Schedule& s, bool fw, bool tr, double* X, const double* B) {
        for (size_t l = 0; l + 1 < s.off.size(); l++) {
            const uint32_t n = s.off[l + 1] - s.off[l];
            hipLaunchKernelGGL(subst_kernel, dim3(nblk(n)), dim3(kT), 0, stream(), f_, nc_, s.cells + s.off[l], n,
                               fw ? 1 : 0, tr ? 1 : 0, X, B, iD_.p, ano_, ann_);
        }
    }
    // precondition_ (solve.cpp:203-218)
    void precondition(const double* R, double* Z, bool tr) {
        using namespace Controls;
        if (Preconditioner == NOPR) {
            copy(Z, R);
        } else if (Preconditioner == DIAG) {
            hipLaunchKernelGGL(diag_kernel, dim3(nblk(nbf_)), dim3(kT), 0, stream(), nc_, Z, R, iD_.p, nbf_);
        } else if (Solver == PCG) {
            const Schedules& s = level_schedules();
            substitute(s.fw, true, tr, Z, R);
            hipLaunchKernelGGL(scale_kernel, dim3(nblk(nall_)), dim3(kT), 0, stream(), nc_, Z, D_.p, nall_);
            substitute(s.bw, false, tr, Z, Z);
        }
    }
    void setup_diagonal();
};

// D = ap, iD = 1 / ap; for symmetric systems SSOR scales them by the
// relaxation factor and DILU takes the incomplete-factorisation diagonal
// (solve.cpp:283-353), in the reference's cell order
void SolverT::setup_diagonal() {
    using namespace Controls;
    std::vector<double> ap(nall_), D(nall_), iD(nall_);
    download(ap.data(), ap_, nall_ * sizeof(double));
    for (uint64_t i = 0; i < nall_; i++) {
        D[i] = ap[i];
        iD[i] = Scalar(1) / ap[i];
    }
    if (Solver == PCG && symmetric()) {
        if (Preconditioner == SSOR) {
            for (uint64_t i = 0; i < nall_; i++) {
                iD[i] *= SOR_omega;
                D[i] *= (2.0 / SOR_omega - 1.0);
            }
        } else if (Preconditioner == DILU) {
            const size_t nfn = m_.FO_h.size();
            std::vector<double> ano(nfn), ann(nfn);
            download(ano.data(), ano_, nfn * sizeof(double));
            download(ann.data(), ann_, nfn * sizeof(double));
            for (uint32_t ci = 0; ci < m_.nbcs; ci++)
                for (uint32_t e = m_.ent_off_h[ci]; e < m_.ent_off_h[ci + 1]; e++) {
                    const uint64_t v = m_.ent_h[e], k = v >> 1;
                    const uint32_t c1 = m_.FO_h[k], c2 = m_.FN_h[k];
                    if (!(v & 1)) {
                        if (c2 > c1 && c2 < nall_) D[c2] -= (ano[k] * ann[k] * iD[c1]);
                    } else {
                        if (c1 > c2) D[c1] -= (ano[k] * ann[k] * iD[c2]);
                    }
                }
            for (uint64_t i = 0; i < nall_; i++) iD[i] = Scalar(1) / D[i];
        }
    }
    upload(D_.p, D.data(), nall_ * sizeof(double));
    upload(iD_.p, iD.data(), nall_ * sizeof(double));
}

void SolverT::run() {
    using namespace Controls;
    const bool pcg = Solver == PCG;
    const bool bicg = pcg && !symmetric();
    if (bicg) {
        r1_.reset(new Buf(nc_ * (nall_ + 1)));
        p1_.reset(new Buf(nc_ * (nall_ + 1)));
        AP1_.reset(new Buf(nc_ * (nall_ + 1)));
    }
    setup_diagonal();
    double* r = r_.p;
    double* p = p_.p;
    double* AP = AP_.p;
    Lanes o_rr{}, oo_rr{}, alpha{}, malpha{}, beta{};
    // CALC_RESID (solve.cpp:248-272): r = Su - A cF, ghosts cleared
    spmv(cF_, AP, false);
    // the reference forms Su - mul(M, cF) element-wise; Su + (A cF)(-1) is
    // the same IEEE value
    axpy(r, Su_, AP, Lanes{{-1, -1, -1, -1, -1, -1, -1, -1, -1}});
    check(hipGetLastError(), "residual");
    zero_ghosts(r);
    check(hipMemsetAsync(AP, 0, (size_t)nc_ * nall_ * sizeof(double), stream()), "solver buffer");
    precondition(r, AP, false);
    zero_ghosts(AP);
    double res = residual(AP);
    if (pcg) {
        o_rr = dot(r, AP);
        copy(p, AP);
        if (bicg) {
            copy(r1_->p, r);
            copy(p1_->p, p);
        }
    }
    const double ires = res;
    int iterations = 0;
    while (iterations < (int)max_iterations) {
        iterations++;
        if (!pcg) {
            copy(p, cF_);
            if (Solver == JACOBI) {
                hipLaunchKernelGGL(rhs_kernel, dim3(nblk(nall_)), dim3(kT), 0, stream(), f_, nc_, cF_, Su_, ann_, ano_,
                                   AP, nall_, nbf_);
                hipLaunchKernelGGL(jacobi_kernel, dim3(nblk(nall_)), dim3(kT), 0, stream(), nc_, cF_, iD_.p, AP,
                                   nall_);
            } else {
                const Schedules& s = level_schedules();
                for (size_t l = 0; l + 1 < s.sor.off.size(); l++) {
                    const uint32_t n = s.sor.off[l + 1] - s.sor.off[l];
                    hipLaunchKernelGGL(sweep_kernel, dim3(nblk(n)), dim3(kT), 0, stream(), f_, nc_,
                                       s.sor.cells + s.sor.off[l], n, cF_, Su_, iD_.p, ano_, ann_, SOR_omega);
                }
            }
            hipLaunchKernelGGL(diff_kernel, dim3(nblk(nbf_)), dim3(kT), 0, stream(), nc_, AP, cF_, p, nbf_);
        } else if (!bicg) {
            spmv(p, AP, false);
            oo_rr = dot(p, AP);
            for (int c = 0; c < nc_; c++) {
                alpha.v[c] = sdiv(o_rr.v[c], oo_rr.v[c]);
                malpha.v[c] = -alpha.v[c];
            }
            axpy(cF_, cF_, p, alpha);
            axpy(r, r, AP, malpha);
            precondition(r, AP, false);
            oo_rr = o_rr;
            o_rr = dot(r, AP);
            for (int c = 0; c < nc_; c++) beta.v[c] = sdiv(o_rr.v[c], oo_rr.v[c]);
            axpy(p, AP, p, beta);
        } else {
            double* r1 = r1_->p;
            double* p1 = p1_->p;
            double* AP1 = AP1_->p;
            spmv(p, AP, false);
            spmv(p1, AP1, true);
            oo_rr = dot(p1, AP);
            for (int c = 0; c < nc_; c++) {
                alpha.v[c] = sdiv(o_rr.v[c], oo_rr.v[c]);
                malpha.v[c] = -alpha.v[c];
            }
            axpy(cF_, cF_, p, alpha);
            axpy(r, r, AP, malpha);
            axpy(r1, r1, AP1, malpha);
            precondition(r, AP, false);
            precondition(r1, AP1, true);
            oo_rr = o_rr;
            o_rr = dot(r1, AP);
            for (int c = 0; c < nc_; c++) beta.v[c] = sdiv(o_rr.v[c], oo_rr.v[c]);
            axpy(p, AP, p, beta);
            axpy(p1, AP1, p1, beta);
        }
        check(hipGetLastError(), "solver iteration");
        res = residual(AP);
        if (res <= tolerance || iterations == (int)max_iterations) break;
    }
    if (MP::printOn) {
        std::string info = symmetric() ? "SYMM-" : "ASYM-";
        info += (flags_ & 2) ? "DIAG-" : "FULL-";
        if (Solver == JACOBI) info += "JAC :";
        else if (Solver == SOR) info += "SOR :";
        else if (Preconditioner == NOPR) info += "NONE-PCG :";
        else if (Preconditioner == DIAG) info += "DIAG-PCG :";
        else if (Preconditioner == SSOR) info += "SSOR-PCG :";
        else info += "DILU-PCG :";
        MP::printH("%sIterations %d Initial Residual %.5e Final Residual %.5e\n", info.c_str(), iterations, ires, res);
    }
}

}  // namespace

void solve_general(const Scalar* ap, const Scalar* ano, const Scalar* ann, const Scalar* Su, Scalar* x, int nc,
                   int flags) {
    if (dg()) fail("Solve: implicit DG systems are not part of this build");
    if (nc < 1 || nc > kMaxNc) fail("Solve: unsupported value type");
    SolverT s(ap, ano, ann, Su, x, nc, flags);
    s.run();
}

}  // namespace nsem_field
