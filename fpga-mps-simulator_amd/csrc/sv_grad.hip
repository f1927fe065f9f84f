// Statevector kernels for adjoint differentiation on gfx950.
//
//  * sv_pauli_inner: c_k = <lam| P_k |psi> for one-qubit Paulis P_k in {X, Y, Z} on
//    qubit q_k, many k per launch (blockIdx.y = k).  With R(theta) = exp(-i theta P / 2)
//    the adjoint gradient of <psi|H|psi> is Im(c_k) at the gate's position.
//  * sv_tfim_apply: out = H psi for H = -J sum Z_i Z_{i+1} - h sum X_i (open chain).
//  * sv_inner: <a|b>.
// Reductions use a fixed grid and fixed summation order (deterministic results).
#include "qsim_internal.h"

namespace qs {

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;

__device__ __forceinline__ double2 cjmul(double2 a, double2 b) {   // conj(a) * b
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.x, b.y, -a.y * b.x));
}

template <int W>
__device__ __forceinline__ void block_reduce2(double& x, double& y, double* red) {
    red[threadIdx.x] = x;
    red[W + threadIdx.x] = y;
    __syncthreads();
    for (int w = W / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x] += red[threadIdx.x + w];
            red[W + threadIdx.x] += red[W + threadIdx.x + w];
        }
        __syncthreads();
    }
    x = red[0];
    y = red[W];
}

// Pair loop over (i0, i1 = i0 | 2^q).
__global__ void __launch_bounds__(kRedThreads)
sv_pauli_inner_kernel(const double2* __restrict__ lam, const double2* __restrict__ psi, int nq,
                      const int32_t* __restrict__ kinds, const int32_t* __restrict__ qubits,
                      double2* __restrict__ part) {
    __shared__ double red[2 * kRedThreads];
    const int k = blockIdx.y;
    const int kind = kinds[k], q = qubits[k];
    const uint64_t half = 1ull << (nq - 1);
    const uint64_t lo = (1ull << q) - 1;
    double sr = 0.0, si = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < half; p += stride) {
        const uint64_t i0 = ((p & ~lo) << 1) | (p & lo);
        const uint64_t i1 = i0 | (1ull << q);
        const double2 l0 = lam[i0], l1 = lam[i1], p0 = psi[i0], p1 = psi[i1];
        double2 t;
        if (kind == 2) {            // Z: conj(l0) p0 - conj(l1) p1
            const double2 a = cjmul(l0, p0), b = cjmul(l1, p1);
            t = make_double2(a.x - b.x, a.y - b.y);
        } else if (kind == 0) {     // X: conj(l0) p1 + conj(l1) p0
            const double2 a = cjmul(l0, p1), b = cjmul(l1, p0);
            t = make_double2(a.x + b.x, a.y + b.y);
        } else {                    // Y: conj(l0) (-i p1) + conj(l1) (i p0)
            const double2 a = cjmul(l0, p1), b = cjmul(l1, p0);
            // -i*a + i*b = (a.y - b.y) + i (b.x - a.x)
            t = make_double2(a.y - b.y, b.x - a.x);
        }
        sr += t.x;
        si += t.y;
    }
    block_reduce2<kRedThreads>(sr, si, red);
    if (threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = make_double2(sr, si);
}

__global__ void __launch_bounds__(kRedThreads)
sv_inner_kernel(const double2* __restrict__ a, const double2* __restrict__ b, uint64_t n,
                double2* __restrict__ part) {
    __shared__ double red[2 * kRedThreads];
    double sr = 0.0, si = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double2 t = cjmul(a[i], b[i]);
        sr += t.x;
        si += t.y;
    }
    block_reduce2<kRedThreads>(sr, si, red);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr, si);
}

__global__ void sv_tfim_kernel(const double2* __restrict__ psi, double2* __restrict__ out, int nq,
                               double J, double h) {
    const uint64_t n = 1ull << nq;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += stride) {
        // sum_i z_i z_{i+1} with z = 1 - 2 bit: (nq - 1) - 2 * popcount(x ^ (x >> 1)) over pairs
        const uint64_t pairs = (x ^ (x >> 1)) & ((n >> 1) - 1);
        const double zz = (double)(nq - 1) - 2.0 * (double)__popcll(pairs);
        const double2 v = psi[x];
        double2 acc = make_double2(-J * zz * v.x, -J * zz * v.y);
        for (int i = 0; i < nq; ++i) {
            const double2 w = psi[x ^ (1ull << i)];
            acc.x -= h * w.x;
            acc.y -= h * w.y;
        }
        out[x] = acc;
    }
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_sv_pauli_inner(const void* lam, const void* psi, int nq, const int32_t* kinds,
                      const int32_t* qubits, int count, double* out_host, void* stream) {
    clear_error();
    QS_REQUIRE(lam && psi && kinds && qubits && out_host, QS_E_ARG, "null pointer");
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    QS_REQUIRE(count >= 0 && count <= 65535, QS_E_RANGE, "count %d outside [0, 65535]", count);
    if (count == 0) return QS_OK;
    for (int k = 0; k < count; ++k) {
        QS_REQUIRE(kinds[k] >= 0 && kinds[k] <= 2, QS_E_ARG, "entry %d: Pauli kind %d not in "
                   "{0=X, 1=Y, 2=Z}", k, kinds[k]);
        QS_REQUIRE(qubits[k] >= 0 && qubits[k] < nq, QS_E_ARG, "entry %d: qubit %d out of range",
                   k, qubits[k]);
    }
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const uint64_t half = 1ull << (nq - 1);
    int blocks = (int)std::min<uint64_t>(kRedBlocks, (half + kRedThreads - 1) / kRedThreads);
    if (blocks < 1) blocks = 1;
    void* dev = nullptr;
    const size_t tb = sv_blob_offset((size_t)count * 2 * sizeof(int32_t));
    const size_t pb = (size_t)count * blocks * sizeof(double2);
    QS_HIP(hipMallocAsync(&dev, tb + pb, hs));
    std::vector<int32_t> tab(2 * (size_t)count);
    for (int k = 0; k < count; ++k) {
        tab[k] = kinds[k];
        tab[count + k] = qubits[k];
    }
    int32_t* dk = static_cast<int32_t*>(dev);
    double2* part = reinterpret_cast<double2*>(static_cast<char*>(dev) + tb);
    QS_HIP(hipMemcpyAsync(dk, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice, hs));
    hipLaunchKernelGGL(sv_pauli_inner_kernel, dim3(blocks, count), dim3(kRedThreads), 0, hs,
                       static_cast<const double2*>(lam), static_cast<const double2*>(psi), nq, dk,
                       dk + count, part);
    QS_HIP(hipGetLastError());
    std::vector<double2> h((size_t)count * blocks);
    QS_HIP(hipMemcpyAsync(h.data(), part, pb, hipMemcpyDeviceToHost, hs));
    QS_HIP(hipFreeAsync(dev, hs));
    QS_HIP(hipStreamSynchronize(hs));
    for (int k = 0; k < count; ++k) {
        double sr = 0.0, si = 0.0;
        for (int b = 0; b < blocks; ++b) {
            sr += h[(size_t)k * blocks + b].x;
            si += h[(size_t)k * blocks + b].y;
        }
        out_host[2 * k] = sr;
        out_host[2 * k + 1] = si;
    }
    return QS_OK;
}

int qs_sv_inner(const void* a, const void* b, int nq, double* out_host, void* stream) {
    clear_error();
    QS_REQUIRE(a && b && out_host, QS_E_ARG, "null pointer");
    QS_REQUIRE(nq >= 1 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [1, %d]", nq,
               SV_MAX_QUBITS);
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const uint64_t n = 1ull << nq;
    int blocks = (int)std::min<uint64_t>(kRedBlocks, (n + kRedThreads - 1) / kRedThreads);
    double2* part = nullptr;
    QS_HIP(hipMallocAsync(reinterpret_cast<void**>(&part), blocks * sizeof(double2), hs));
    hipLaunchKernelGGL(sv_inner_kernel, dim3(blocks), dim3(kRedThreads), 0, hs,
                       static_cast<const double2*>(a), static_cast<const double2*>(b), n, part);
    QS_HIP(hipGetLastError());
    std::vector<double2> h(blocks);
    QS_HIP(hipMemcpyAsync(h.data(), part, blocks * sizeof(double2), hipMemcpyDeviceToHost, hs));
    QS_HIP(hipFreeAsync(part, hs));
    QS_HIP(hipStreamSynchronize(hs));
    double sr = 0.0, si = 0.0;
    for (const double2& v : h) {
        sr += v.x;
        si += v.y;
    }
    out_host[0] = sr;
    out_host[1] = si;
    return QS_OK;
}

int qs_sv_apply_tfim(const void* psi, void* out, int nq, double J, double h, void* stream) {
    clear_error();
    QS_REQUIRE(psi && out && psi != out, QS_E_ARG, "need distinct non-null psi and out");
    QS_REQUIRE(nq >= 2 && nq <= SV_MAX_QUBITS, QS_E_RANGE, "nq=%d outside [2, %d]", nq,
               SV_MAX_QUBITS);
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const uint64_t n = 1ull << nq;
    const int blocks = (int)std::min<uint64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(sv_tfim_kernel, dim3(blocks), dim3(256), 0, hs,
                       static_cast<const double2*>(psi), static_cast<double2*>(out), nq, J, h);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

}  // extern "C"
