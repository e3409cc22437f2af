// tlw_noalloc.hpp -- tlw_max_dt_noalloc: Trixi's max_dt on a TreeMesh state
// without a context and without any device allocation.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) pass the state, the
// element inverse Jacobians and a device pointer for the result.  The
// per-workgroup partial maxima live in `scratch`, a fixed device array of this
// file; each GPU the library runs on holds its own copy until the process
// ends.  Two launches on the caller's stream, no host synchronisation:
//   k_na_speed   one lane per GLL node, EB = 128 / (N+1)^2 elements per
//                workgroup: per element J⁻¹ (max λ1 + max λ2), λ_d = |v_d| + c,
//                the workgroup's maximum (or -1 if a node has ρ <= 0 or p <= 0)
//                into scratch[workgroup]
//   k_na_final   one workgroup: max over the partials, max_dt = 2 / (n M)
//                (NaN if any partial flagged a non-admissible node)
// The arithmetic is that of tlw_max_dt (the same node speeds, per-element
// products and order-free maxima), so both give the same bits.
//
// Every workgroup of k_na_speed writes its own partial before k_na_final
// reads them, so nothing has to start at zero.  Calls that overlap on one GPU
// (different streams or host threads) share `scratch` and race.
#pragma once

#include <mutex>
#include <string>

namespace tlw_noalloc {

using namespace tlw;

// 2 MiB of partial maxima: ceil(n_elements / EB) must not exceed this
// (polydeg 4: EB = 5, up to 1 310 720 elements; polydeg 3: EB = 8, 2 097 152).
constexpr int kScratchDoubles = 1 << 18;
constexpr int kFinalThreads = 1024;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

template <int N>
__global__ __launch_bounds__(kSurfBlock) void k_na_speed(const double* __restrict__ u,
                                                       const double* __restrict__ inv_jac, int E, double gamma,
                                                       double* __restrict__ partial) {
    using G_ = SGeo<N>;
    constexpr int NN = G_::NN, EB = G_::EB;
    __shared__ double s_l1[EB * NN], s_l2[EB * NN], s_m[EB];
    __shared__ int s_bad;
    const int tid = threadIdx.x;
    const int le = tid / NN, node = tid - le * NN;
    const int e0 = blockIdx.x * EB;
    const int nel = min(EB, E - e0);
    const bool act = le < nel;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    if (act) {
        const int e = e0 + le;
        double val[4];
        load4(u + ((size_t)e * NN + node) * 4, val);
        double l1, l2;
        bool bad;
        node_speeds(val, gamma, l1, l2, bad);
        s_l1[le * NN + node] = l1;
        s_l2[le * NN + node] = l2;
        if (bad) s_bad = 1;
    }
    __syncthreads();
    if (act && node == 0) {
        double m1 = 0.0, m2 = 0.0;
        for (int k = 0; k < NN; ++k) {
            m1 = dmax(m1, s_l1[le * NN + k]);
            m2 = dmax(m2, s_l2[le * NN + k]);
        }
        s_m[le] = inv_jac[e0 + le] * (m1 + m2);
    }
    __syncthreads();
    if (tid == 0) {
        double bm = 0.0;
        for (int k = 0; k < nel; ++k) bm = dmax(bm, s_m[k]);
        partial[blockIdx.x] = s_bad ? -1.0 : bm;
    }
}

template <int N>
__global__ __launch_bounds__(kFinalThreads) void k_na_final(const double* __restrict__ partial, int G,
                                                          double* __restrict__ out) {
    constexpr int n = N + 1;
    __shared__ double s[kFinalThreads];
    __shared__ int s_bad;
    const int t = threadIdx.x;
    if (t == 0) s_bad = 0;
    __syncthreads();
    double m = 0.0;
    bool bad = false;
    for (int k = t; k < G; k += kFinalThreads) {
        const double p = partial[k];
        bad = bad || p < 0.0;
        m = dmax(m, p);
    }
    if (bad) s_bad = 1;
    s[t] = m;
    __syncthreads();
    for (int w = kFinalThreads / 2; w > 0; w >>= 1) {
        if (t < w) s[t] = dmax(s[t], s[t + w]);
        __syncthreads();
    }
    if (t == 0) {
        const double tiny = 4.9406564584124654e-324;  // nextfloat(0.0), Trixi's start value
        out[0] = s_bad ? __builtin_nan("") : 2.0 / (n * dmax(tiny, s[0]));
    }
}

// Device address of `scratch` on the current device, looked up once per
// device with hipGetSymbolAddress and kept: host code never passes the
// array's own name to a kernel (on the host it names a stand-in, not device
// memory).
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

static int scratch_on_current_device(double** out, std::string& err) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return TLW_EHIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        err = "device ordinal out of range";
        return TLW_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return TLW_EHIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return TLW_OK;
}

template <int N> int launch(const double* u, const double* inv_jac, int E, double gamma, double* out, hipStream_t s,
                            std::string& err) {
    constexpr int EB = SGeo<N>::EB;
    const long long G = ((long long)E + EB - 1) / EB;
    if (G > kScratchDoubles) {
        err = "n_elements exceeds the scratch capacity of tlw_max_dt_noalloc (" +
              std::to_string((long long)kScratchDoubles * EB) + " elements at polydeg " + std::to_string(N) + ")";
        return TLW_EINVAL;
    }
    double* part = nullptr;
    int rc = scratch_on_current_device(&part, err);
    if (rc) return rc;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_na_speed<N>), dim3((unsigned)G), dim3(kSurfBlock), 0, s, u, inv_jac, E, gamma,
                       part);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_na_final<N>), dim3(1), dim3(kFinalThreads), 0, s, part, (int)G, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("kernel launch: ") + hipGetErrorString(e);
        return TLW_EHIP;
    }
    return TLW_OK;
}

}  // namespace tlw_noalloc
