// Measurement shots of a device-resident statevector on gfx950, as the reference's driver
// draws them: probs = |a|^2 / sum |a|^2, indices = np.random.choice(2^n, shots, p=probs)
// (fpga_simulator_v07.py:592-609).  numpy's legacy choice forms cdf = cumsum(probs),
// divides it by its last entry, draws one uniform u per shot from the generator and
// returns searchsorted(cdf, u, 'right'): the first index whose cumulative probability
// exceeds u.  Here the caller draws the same uniforms on the host (so a seeded global
// generator gives the reference's stream) and the device does the rest:
//
//  * block sums: 2^n / kBlock blocks of kBlock amplitudes, each summed in a fixed order;
//  * prefix: one workgroup turns the block sums into exclusive prefixes pre[b] and the
//    total pre[nb] (fixed order, deterministic);
//  * search: one wavefront per shot finds the first block with u total < pre[b + 1] by
//    bisection, then the amplitude inside the block from the lanes' partial sums.
//    Amplitudes of probability 0 are never returned (first index whose inclusive prefix
//    exceeds the target); when rounding puts the target past the block's own sum, the
//    block's last amplitude of nonzero probability is returned.
//
// Cumulative sums are rounded differently from numpy's sequential cumsum, so an index
// can differ from the host's only when u falls within rounding distance (about 1e-16
// relative) of a cumulative-probability boundary.
#include "qsim_internal.h"

namespace qs {

namespace {

constexpr int kBlock = 4096;              // amplitudes per prefix block
constexpr int kSumThreads = 256;          // block-sum kernel: 16 amplitudes per thread
constexpr int kScanThreads = 1024;
constexpr int kWave = 64;
constexpr int kShotsPerWg = 4;
constexpr int kPerLane = kBlock / kWave;  // 64 amplitudes per lane in the search

__device__ __forceinline__ double prob(const cd2* s, uint64_t i) {
    const cd2 a = s[i];
    return fma(a.x, a.x, a.y * a.y);
}

// sums[b] = sum of |a|^2 over block b (thread t adds amplitudes t, t + 256, ...; the
// 256 partial sums are added pairwise in LDS).
__global__ void __launch_bounds__(kSumThreads)
sv_block_sums_kernel(const cd2* __restrict__ state, uint64_t n, double* __restrict__ sums) {
    __shared__ double part[kSumThreads];
    const uint64_t b0 = (uint64_t)blockIdx.x * kBlock;
    const uint64_t len = n - b0 < (uint64_t)kBlock ? n - b0 : (uint64_t)kBlock;
    double v = 0.0;
    for (uint64_t i = threadIdx.x; i < len; i += kSumThreads) v += prob(state, b0 + i);
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = kSumThreads / 2; w >= 1; w /= 2) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// In place: pre[b] = sum_{c < b} sums[c] for b < nb, and pre[nb] = the total.  Thread t
// owns a contiguous chunk of blocks; the chunk totals are scanned by thread 0.
__global__ void __launch_bounds__(kScanThreads)
sv_prefix_kernel(double* __restrict__ pre, int nb) {
    __shared__ double chunk[kScanThreads];
    const int per = (nb + kScanThreads - 1) / kScanThreads;
    const int c0 = threadIdx.x * per;
    const int c1 = c0 + per < nb ? c0 + per : nb;
    double s = 0.0;
    for (int c = c0; c < c1; ++c) s += pre[c];
    chunk[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < kScanThreads; ++t) {
            const double x = chunk[t];
            chunk[t] = run;
            run += x;
        }
        pre[nb] = run;
    }
    __syncthreads();
    double run = chunk[threadIdx.x];
    for (int c = c0; c < c1; ++c) {
        const double x = pre[c];
        pre[c] = run;
        run += x;
    }
}

// One wavefront per shot.
__global__ void __launch_bounds__(kWave * kShotsPerWg)
sv_sample_kernel(const cd2* __restrict__ state, uint64_t n, const double* __restrict__ pre,
                 int nb, const double* __restrict__ uniforms, int shots,
                 int64_t* __restrict__ out) {
    const int lane = threadIdx.x % kWave;
    const int shot = blockIdx.x * kShotsPerWg + threadIdx.x / kWave;
    if (shot >= shots) return;                       // the whole wavefront leaves
    const double total = pre[nb];
    const double target = uniforms[shot] * total;
    // block: the first b whose cumulative sum pre[b + 1] exceeds the target (blocks of
    // probability 0 are passed over; pre[nb] = total > target since u < 1)
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (pre[mid + 1] > target) hi = mid; else lo = mid + 1;
    }
    const int b = lo;
    const uint64_t b0 = (uint64_t)b * kBlock;
    const uint64_t len = n - b0 < (uint64_t)kBlock ? n - b0 : (uint64_t)kBlock;
    const double t = target - pre[b];
    // lane l sums amplitudes [64 l, 64 l + 64) of the block in order
    const uint64_t l0 = (uint64_t)lane * kPerLane;
    double ls = 0.0;
    for (int k = 0; k < kPerLane; ++k)
        if (l0 + k < len) ls += prob(state, b0 + l0 + k);
    // inclusive prefix over lanes, in lane order
    double inc = ls;
    for (int o = 1; o < kWave; o *= 2) {
        const double y = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += y;
    }
    const double exc = inc - ls;
    // the first lane whose inclusive prefix exceeds t, else the last lane holding
    // probability
    const uint64_t over = __ballot(inc > t && ls > 0.0);
    const uint64_t nz = __ballot(ls > 0.0);
    int L;
    bool past = false;
    if (over) {
        L = __ffsll((unsigned long long)over) - 1;
    } else {
        L = nz ? 63 - __clzll((unsigned long long)nz) : 0;
        past = true;
    }
    if (lane == L) {
        int64_t idx = -1;
        double run = exc;
        int64_t last = -1;
        for (int k = 0; k < kPerLane && l0 + k < len; ++k) {
            const double p = prob(state, b0 + l0 + k);
            if (p > 0.0) last = (int64_t)(b0 + l0 + k);
            run += p;
            if (!past && idx < 0 && p > 0.0 && run > t) idx = (int64_t)(b0 + l0 + k);
        }
        out[shot] = idx >= 0 ? idx : last;
    }
}

}  // namespace

size_t sv_sample_workspace_bytes(int nq) {
    const uint64_t n = 1ull << nq;
    const uint64_t nb = (n + kBlock - 1) / kBlock;
    return (size_t)(nb + 1) * sizeof(double);
}

int sv_sample(const double2* state, int nq, const double* uniforms, int shots, int64_t* out,
              double* ws, size_t ws_bytes, hipStream_t stream) {
    const uint64_t n = 1ull << nq;
    const int nb = (int)((n + kBlock - 1) / kBlock);
    QS_REQUIRE(ws_bytes >= sv_sample_workspace_bytes(nq), QS_E_ARG,
               "sampling workspace of %zu bytes, %zu needed", ws_bytes,
               sv_sample_workspace_bytes(nq));
    const cd2* s = reinterpret_cast<const cd2*>(state);
    hipLaunchKernelGGL(sv_block_sums_kernel, dim3(nb), dim3(kSumThreads), 0, stream, s, n, ws);
    QS_HIP(hipGetLastError());
    hipLaunchKernelGGL(sv_prefix_kernel, dim3(1), dim3(kScanThreads), 0, stream, ws, nb);
    QS_HIP(hipGetLastError());
    double total = 0.0;
    QS_HIP(hipMemcpyAsync(&total, ws + nb, sizeof(double), hipMemcpyDeviceToHost, stream));
    QS_HIP(hipStreamSynchronize(stream));
    QS_REQUIRE(total > 0.0 && std::isfinite(total), QS_E_ARG,
               "probabilities do not sum to 1 (sum of |a|^2 = %g)", total);
    const int wgs = (shots + kShotsPerWg - 1) / kShotsPerWg;
    hipLaunchKernelGGL(sv_sample_kernel, dim3(wgs), dim3(kWave * kShotsPerWg), 0, stream, s, n,
                       ws, nb, uniforms, shots, out);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

}  // namespace qs
