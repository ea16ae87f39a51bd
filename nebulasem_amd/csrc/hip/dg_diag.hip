// Per-step diagnostics of the Euler solver on the device: Courant number
// max/min/avg (Mesh::calc_courant, field.cpp:440-448) and the mass, energy
// and volume sums of euler.cpp:261-283 over the interior nodes.
//
// Deterministic: a fixed grid of blocks, each thread accumulating its
// grid-stride nodes in order, block partials to slots, then one block
// reducing the slots in a fixed order (after a kernel boundary, so every
// slot is visible across XCDs).
#include <mutex>

#include "dd_math.hpp"
#include "dg_context.hpp"
#include "dg_kernels.hpp"

namespace nsem {

namespace {

constexpr int kDiagThreads = 256;

struct Acc {
    double cmax, cmin, csum, mass, energy, vol;
};

__device__ __forceinline__ void combine(Acc& a, const Acc& b) {
    a.cmax = fmax(a.cmax, b.cmax);
    a.cmin = fmin(a.cmin, b.cmin);
    a.csum += b.csum; a.mass += b.mass; a.energy += b.energy; a.vol += b.vol;
}

__device__ __forceinline__ Acc shfl_down(const Acc& a, int off) {
    Acc r;
    r.cmax = __shfl_down(a.cmax, off, 64); r.cmin = __shfl_down(a.cmin, off, 64);
    r.csum = __shfl_down(a.csum, off, 64); r.mass = __shfl_down(a.mass, off, 64);
    r.energy = __shfl_down(a.energy, off, 64); r.vol = __shfl_down(a.vol, off, 64);
    return r;
}

__device__ Acc block_reduce(Acc a) {
    __shared__ Acc part[kDiagThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        Acc o = shfl_down(a, off);
        combine(a, o);
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) part[w] = a;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < kDiagThreads / 64; k++) combine(a, part[k]);
    return a;
}

__global__ void __launch_bounds__(kDiagThreads) diag_partial_kernel(DgDev d, DiagState s, const double* gh,
                                                                    double* slots) {
    Acc a{-1e30, 1e30, 0, 0, 0, 0};
    const double R = d.R, cv = d.cp - d.R;
    const uint64_t stride = (uint64_t)gridDim.x * kDiagThreads;
    for (uint64_t n = (uint64_t)blockIdx.x * kDiagThreads + threadIdx.x; n < d.nb; n += stride) {
        double rho, u[3], th, pp;
        if (s.mol) {
            const double rr = d.rho_ref[n];
            rho = rr + s.q[0][n];
            for (int c = 0; c < 3; c++) u[c] = s.q[1 + c][n] / rho;
            th = (s.q[4][n] + rr * d.T0) / rho;
            pp = d.p_ref[n] * expm1(d.gamma * log1p(s.q[4][n] / (rr * d.T0)));
        } else {
            rho = s.rho[n];
            u[0] = s.ux[n]; u[1] = s.uy[n]; u[2] = s.uz[n];
            th = s.T[n] + d.T0;
            pp = s.p[n];
        }
        const double cvn = d.cV[n];
        const double uu = dot3r(u[0], u[1], u[2], u[0], u[1], u[2]);
        const double c = sqrt(uu) * d.dt / dd::pow_dd(cvn, 1.0 / 3);
        a.cmax = fmax(a.cmax, fabs(c));
        a.cmin = fmin(a.cmin, fabs(c));
        a.csum += c;
        a.mass += rho * cvn;
        double sf = gh[n] + 0.5 * uu + dd::pow_dd((pp + d.p_ref[n]) / d.P0, R / d.cp) * th * cv;
        a.energy += rho * cvn * sf;
        a.vol += cvn;
    }
    a = block_reduce(a);
    if (threadIdx.x == 0) {
        double* o = slots + (size_t)blockIdx.x * 6;
        o[0] = a.cmax; o[1] = a.cmin; o[2] = a.csum; o[3] = a.mass; o[4] = a.energy; o[5] = a.vol;
    }
}

__global__ void __launch_bounds__(kDiagThreads) diag_final_kernel(const double* slots, unsigned nblocks,
                                                                  uint64_t nb, double* out) {
    Acc a{-1e30, 1e30, 0, 0, 0, 0};
    for (unsigned b = threadIdx.x; b < nblocks; b += kDiagThreads) {
        const double* o = slots + (size_t)b * 6;
        Acc v{o[0], o[1], o[2], o[3], o[4], o[5]};
        combine(a, v);
    }
    a = block_reduce(a);
    if (threadIdx.x == 0) {
        out[0] = a.cmax; out[1] = a.cmin; out[2] = a.csum / (double)nb;
        out[3] = a.mass; out[4] = a.energy; out[5] = a.vol;
    }
}

}  // namespace

hipError_t launch_diagnostics(const DgDev& d, const DiagState& s, const double* gh, double* slots,
                              double* out, hipStream_t st) {
    hipLaunchKernelGGL(diag_partial_kernel, dim3(kDiagBlocks), dim3(kDiagThreads), 0, st, d, s, gh, slots);
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(kDiagThreads), 0, st, slots, (unsigned)kDiagBlocks,
                       d.nb, out);
    return hipGetLastError();
}

}  // namespace nsem

// ---------------------------------------------------------------------------
// Scratch of nsem_dg_diagnostics_noalloc: the kDiagBlocks block partials (6
// doubles each) followed by the 6 results, 49,200 bytes, the same for every
// problem size.  Each device holds one copy until the process ends; every
// call on a device shares it (each call rewrites all of it before reading),
// so calls that overlap on one device race.
constexpr size_t kDiagScratchDoubles = (size_t)nsem::kDiagBlocks * 6 + 6;
static __device__ __align__(256) double scratch[kDiagScratchDoubles];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is not a device address).
hipError_t scratch_address(int device, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int nsem_dg_diagnostics_noalloc(nsem_dg* d, double* out) {
    if (!d || !out) return nsem_fail(NSEM_EINVAL, "null argument");
    static_assert(kDiagScratchDoubles >= (size_t)nsem::kDiagBlocks * 6 + 6, "diagnostics scratch too small");
    if (d->device < 0 || d->device >= kMaxScratchDevices)
        return nsem_fail(NSEM_EUNSUPPORTED, "device index beyond the scratch table (64 devices)");
    int current = 0;
    HIP_TRY(hipGetDevice(&current));
    if (current != d->device) HIP_TRY(hipSetDevice(d->device));
    double* ptr = nullptr;
    hipError_t e = scratch_address(d->device, &ptr);
    if (e == hipSuccess) {
        double* slots = ptr;
        double* res = ptr + (size_t)nsem::kDiagBlocks * 6;
        e = nsem::launch_diagnostics(d->dev, nsem_dg_diag_state(d), d->gh.as<double>(), slots, res, d->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(out, res, 6 * sizeof(double), hipMemcpyDeviceToHost, d->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    }
    if (current != d->device) (void)hipSetDevice(current);
    if (e != hipSuccess) return nsem_fail(NSEM_EDEVICE, std::string("diagnostics (no allocation): ") + hipGetErrorString(e));
    return NSEM_OK;
}
