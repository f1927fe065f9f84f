// Shot sampling of a device-resident MPS chain on gfx950: right environments and
// sequential conditional sampling (mps_core.py:354-420), all shots in parallel.
//
//  * Right environments R_n = [[1]], R_i = sum_s A_i[:, s, :] R_{i+1} A_i[:, s, :]^H,
//    divided by its trace (when above 1e-300).  One launch per site (the recursion is
//    sequential over sites); within a site, workgroup (ta, tb) computes the 16 x 16 tile
//    a in [16 ta, 16 ta + 16), a' in [16 tb, 16 tb + 16): first T_s = A_s[a rows] R_{i+1}
//    into LDS, then R_i[a, a'] = sum_s sum_c T_s[a, c] conj(A_s[a', c]).  Environments are
//    stored as computed; the next site divides by the trace it reads off the diagonal, and
//    a last launch divides every environment by its own trace, so the stored R_i equal the
//    recursion's normalised ones.  Traces are summed in a fixed order (deterministic).
//  * Sampling: one wavefront per shot.  Per site: Av_s = v A[:, s, :], p_s =
//    Re(Av_s R_{i+1} Av_s^H), clipped at 0 and normalised (0.5 / 0.5 below 1e-15); the
//    outcome is 1 exactly when the shot's uniform u >= p_0 / (p_0 + p_1) as numpy's
//    Generator.choice(2, p=...) decides it; then v = Av_outcome / ||Av_outcome|| (when the
//    norm exceeds 1e-15).  The uniforms come from the caller, one per (shot, site) in
//    shot-major order, which is the order the reference draws them in.
#include <cmath>
#include <vector>

#include "qsim_internal.h"

namespace qs {

namespace {

constexpr int kEnvTile = 16;
constexpr int kMaxEnvChi = 128;
constexpr int kShotLanes = 64;
constexpr int kShotsPerBlock = 4;

__device__ __forceinline__ cd2 cmul(cd2 a, cd2 b) {
    return mk2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ cd2 cmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ cd2 cjmadd(cd2 acc, cd2 a, cd2 b) {   // acc + a * conj(b)
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.y, b.x, fma(-a.x, b.y, acc.y));
    return acc;
}

__device__ __forceinline__ const cd2* site_ptr(const double* sites, int64_t stride, int k) {
    return reinterpret_cast<const cd2*>(sites) + (size_t)k * (size_t)stride;
}

// Sum of the real diagonal of a (dim x dim) environment stored with row stride ld, in
// index order (every thread that calls it gets the same value).
__device__ __forceinline__ double env_trace(const cd2* R, int dim, int ld) {
    double tr = 0.0;
    for (int k = 0; k < dim; ++k) tr += R[(size_t)k * ld + k].x;
    return tr;
}

__global__ void mps_env_init_kernel(cd2* __restrict__ envs, int n, int ld) {
    if (threadIdx.x == 0 && blockIdx.x == 0) envs[(size_t)n * ld * ld] = mk2(1.0, 0.0);
}

__global__ void __launch_bounds__(kEnvTile * kEnvTile)
mps_env_site_kernel(const double* __restrict__ sites, const int32_t* __restrict__ bond,
                    int64_t stride, int site, int ld, cd2* __restrict__ envs) {
    __shared__ cd2 T[2][kEnvTile][kMaxEnvChi];
    __shared__ cd2 Ap[2][kEnvTile][kMaxEnvChi];
    __shared__ double sh_scale;
    const int cl = bond[site], cr = bond[site + 1];
    const int a0 = blockIdx.x * kEnvTile, b0 = blockIdx.y * kEnvTile;
    if (a0 >= cl || b0 >= cl) return;                 // uniform per workgroup
    const cd2* A = site_ptr(sites, stride, site);     // (cl, 2, cr), C order
    const cd2* R = envs + (size_t)(site + 1) * ld * ld;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const double tr = env_trace(R, cr, ld);
        sh_scale = tr > 1e-300 ? 1.0 / tr : 1.0;
    }
    __syncthreads();
    const double scale = sh_scale;
    // T_s[a, c] = sum_b A[a, s, b] R[b, c] (R divided by its trace) for the tile's rows,
    // and the tile's a' rows of A.
    for (int idx = tid; idx < 2 * kEnvTile * cr; idx += kEnvTile * kEnvTile) {
        const int s = idx / (kEnvTile * cr);
        const int rem = idx - s * kEnvTile * cr;
        const int al = rem / cr, c = rem - al * cr;
        const int a = a0 + al, ap = b0 + al;
        cd2 acc = mk2(0.0, 0.0);
        if (a < cl) {
            const cd2* arow = A + ((size_t)a * 2 + s) * cr;
            for (int b = 0; b < cr; ++b) acc = cmadd(acc, arow[b], R[(size_t)b * ld + c]);
        }
        T[s][al][c] = mk2(acc.x * scale, acc.y * scale);
        Ap[s][al][c] = ap < cl ? A[((size_t)ap * 2 + s) * cr + c] : mk2(0.0, 0.0);
    }
    __syncthreads();
    const int al = tid / kEnvTile, bl = tid % kEnvTile;
    const int a = a0 + al, ap = b0 + bl;
    if (a < cl && ap < cl) {
        cd2 acc = mk2(0.0, 0.0);
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < cr; ++c) acc = cjmadd(acc, T[s][al][c], Ap[s][bl][c]);
        envs[(size_t)site * ld * ld + (size_t)a * ld + ap] = acc;
    }
}

// Divides environment i (i < n) by its trace; one workgroup per environment.
__global__ void __launch_bounds__(256)
mps_env_normalise_kernel(const int32_t* __restrict__ bond, int n, int ld,
                         cd2* __restrict__ envs) {
    __shared__ double sh_scale;
    const int i = blockIdx.x;
    if (i >= n) return;
    const int dim = bond[i];
    cd2* R = envs + (size_t)i * ld * ld;
    if (threadIdx.x == 0) {
        const double tr = env_trace(R, dim, ld);
        sh_scale = tr > 1e-300 ? 1.0 / tr : 1.0;
    }
    __syncthreads();
    const double s = sh_scale;
    for (int idx = threadIdx.x; idx < dim * dim; idx += blockDim.x) {
        const int r = idx / dim, c = idx - r * dim;
        const cd2 v = R[(size_t)r * ld + c];
        R[(size_t)r * ld + c] = mk2(v.x * s, v.y * s);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kShotLanes / 2; o >= 1; o /= 2) v += __shfl_xor(v, o, kShotLanes);
    return v;
}

__global__ void __launch_bounds__(kShotLanes * kShotsPerBlock)
mps_sample_kernel(const double* __restrict__ sites, const int32_t* __restrict__ bond,
                  int64_t stride, int n, int ld, const cd2* __restrict__ envs,
                  const double* __restrict__ uniforms, int shots,
                  uint8_t* __restrict__ bits) {
    __shared__ cd2 vbuf[kShotsPerBlock][kMaxEnvChi];
    __shared__ cd2 avbuf[kShotsPerBlock][2][kMaxEnvChi];
    const int w = threadIdx.x / kShotLanes, lane = threadIdx.x % kShotLanes;
    const int shot = blockIdx.x * kShotsPerBlock + w;
    if (shot >= shots) return;                         // whole wavefronts only
    cd2* v = vbuf[w];
    if (lane == 0) v[0] = mk2(1.0, 0.0);
    __builtin_amdgcn_wave_barrier();
    for (int i = 0; i < n; ++i) {
        const int cl = bond[i], cr = bond[i + 1];
        const cd2* A = site_ptr(sites, stride, i);
        const cd2* R = envs + (size_t)(i + 1) * ld * ld;
        double p[2];
        for (int s = 0; s < 2; ++s) {
            // Av_s[k] = sum_j v[j] A[j, s, k]
            for (int k = lane; k < cr; k += kShotLanes) {
                cd2 acc = mk2(0.0, 0.0);
                for (int j = 0; j < cl; ++j) acc = cmadd(acc, v[j], A[((size_t)j * 2 + s) * cr + k]);
                avbuf[w][s][k] = acc;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // p_s = Re sum_k (sum_i Av[i] R[i, k]) conj(Av[k])
            double part = 0.0;
            for (int k = lane; k < cr; k += kShotLanes) {
                cd2 acc = mk2(0.0, 0.0);
                for (int ii = 0; ii < cr; ++ii) acc = cmadd(acc, avbuf[w][s][ii], R[(size_t)ii * ld + k]);
                const cd2 av = avbuf[w][s][k];
                part += acc.x * av.x + acc.y * av.y;
            }
            p[s] = wave_sum(part);
        }
        double p0 = fmax(p[0], 0.0), p1 = fmax(p[1], 0.0);
        const double tot = p0 + p1;
        if (tot < 1e-15) {
            p0 = 0.5;
            p1 = 0.5;
        } else {
            p0 /= tot;
            p1 /= tot;
        }
        const double cdf0 = p0 / (p0 + p1);
        const double u = uniforms[(size_t)shot * n + i];
        const int o = u >= cdf0 ? 1 : 0;
        if (lane == 0) bits[(size_t)shot * n + i] = (uint8_t)o;
        // v = Av_o / ||Av_o||
        double nrm = 0.0;
        for (int k = lane; k < cr; k += kShotLanes) {
            const cd2 x = avbuf[w][o][k];
            nrm += x.x * x.x + x.y * x.y;
        }
        nrm = sqrt(wave_sum(nrm));
        const double inv = nrm > 1e-15 ? 1.0 / nrm : 1.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int k = lane; k < cr; k += kShotLanes) {
            const cd2 x = avbuf[w][o][k];
            v[k] = mk2(x.x * inv, x.y * inv);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// One step of the dense contraction (mps_core.py:433-461): with the state of sites 0..i-1
// as cur[x][l] (x < 2^i, site q on bit q of x, l < chi_l), nxt[x + s 2^i][r] =
// f sum_l cur[x][l] A_i[l][s][r].  One thread per output element; threads of one x read
// the same row of cur.  Step 0 copies site 0 (cur = nullptr).
__global__ void mps_dense_step_kernel(const cd2* __restrict__ sites, int64_t stride, int i,
                                      int cl, int cr, const cd2* __restrict__ cur,
                                      cd2* __restrict__ nxt, double f) {
    const uint64_t total = (2ull << i) * (uint64_t)cr;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= total) return;
    const uint64_t xo = tid / cr;
    const int r = (int)(tid - xo * cr);
    const int sidx = (int)(xo >> i);
    const uint64_t x = xo & ((1ull << i) - 1);
    const cd2* A = sites + (int64_t)i * stride;
    cd2 acc = mk2(0.0, 0.0);
    if (cur == nullptr) {
        acc = A[sidx * cr + r];                      // site 0: chi_l = 1
    } else {
        for (int l = 0; l < cl; ++l) acc = cmadd(acc, cur[x * cl + l], A[(l * 2 + sidx) * cr + r]);
    }
    nxt[tid] = mk2(acc.x * f, acc.y * f);
}

int check_chain(const double* sites, const int32_t* bond, int64_t stride, int n, int chi_max) {
    QS_REQUIRE(sites != nullptr && bond != nullptr, QS_E_ARG, "null sites or bond array");
    QS_REQUIRE(n >= 1, QS_E_ARG, "chain of %d sites", n);
    QS_REQUIRE(chi_max >= 1 && chi_max <= kMaxEnvChi, QS_E_CHI, "chi_max %d outside [1, %d]",
               chi_max, kMaxEnvChi);
    QS_REQUIRE(stride >= (int64_t)2 * chi_max * chi_max, QS_E_ARG,
               "site_stride %lld < 2*chi_max^2", (long long)stride);
    return QS_OK;
}

}  // namespace

}  // namespace qs

using namespace qs;

extern "C" {

size_t qs_mps_env_bytes(int n, int chi_max) {
    if (n < 1 || chi_max < 1) return 0;
    return (size_t)(n + 1) * chi_max * chi_max * sizeof(cd2);
}

int qs_mps_right_envs(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                      int chi_max, double* envs, size_t env_bytes, void* stream) {
    clear_error();
    int st = check_chain(sites, bond, site_stride, n, chi_max);
    if (st) return st;
    QS_REQUIRE(envs != nullptr, QS_E_ARG, "null environment buffer");
    QS_REQUIRE(env_bytes >= qs_mps_env_bytes(n, chi_max), QS_E_RANGE,
               "environment buffer %zu bytes < %zu needed", env_bytes,
               qs_mps_env_bytes(n, chi_max));
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    cd2* E = reinterpret_cast<cd2*>(envs);
    hipLaunchKernelGGL(mps_env_init_kernel, dim3(1), dim3(64), 0, hs, E, n, chi_max);
    QS_HIP(hipGetLastError());
    const int tiles = (chi_max + kEnvTile - 1) / kEnvTile;
    for (int i = n - 1; i >= 0; --i) {
        hipLaunchKernelGGL(mps_env_site_kernel, dim3(tiles, tiles), dim3(kEnvTile * kEnvTile), 0,
                           hs, sites, bond, site_stride, i, chi_max, E);
        QS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(mps_env_normalise_kernel, dim3(n), dim3(256), 0, hs, bond, n, chi_max, E);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

int qs_mps_statevector(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                       int chi_max, double norm_log, double* out, double* work,
                       size_t buffer_bytes, void* stream) {
    clear_error();
    int st = check_chain(sites, bond, site_stride, n, chi_max);
    if (st) return st;
    QS_REQUIRE(n <= 30, QS_E_RANGE, "dense state of %d qubits", n);
    QS_REQUIRE(out != nullptr && work != nullptr, QS_E_ARG, "null output or work buffer");
    const size_t need = ((size_t)1 << n) * sizeof(cd2);
    QS_REQUIRE(buffer_bytes >= need, QS_E_RANGE, "buffers of %zu bytes, %zu needed",
               buffer_bytes, need);
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    std::vector<int32_t> b(n + 1);
    QS_HIP(hipMemcpyAsync(b.data(), bond, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, hs));
    QS_HIP(hipStreamSynchronize(hs));
    QS_REQUIRE(b[0] == 1 && b[n] == 1, QS_E_ARG, "boundary bonds %d, %d (must be 1)", b[0], b[n]);
    for (int i = 0; i < n; ++i) {
        QS_REQUIRE(b[i + 1] >= 1 && b[i + 1] <= chi_max, QS_E_CHI, "bond %d = %d", i + 1,
                   b[i + 1]);
        // every intermediate fits in 2^n amplitudes when bond i+1 <= 2^(n-i-1), which
        // SVD truncation guarantees
        QS_REQUIRE(((uint64_t)2 << i) * (uint64_t)b[i + 1] <= ((uint64_t)1 << n), QS_E_RANGE,
                   "bond %d = %d exceeds the dimension of the sites to its right", i + 1,
                   b[i + 1]);
    }
    const cd2* S = reinterpret_cast<const cd2*>(sites);
    cd2* bufs[2] = {reinterpret_cast<cd2*>(out), reinterpret_cast<cd2*>(work)};
    const double f = std::exp(norm_log);
    const cd2* cur = nullptr;
    for (int i = 0; i < n; ++i) {
        cd2* nxt = bufs[(n - 1 - i) & 1];            // the last step writes `out`
        const uint64_t total = ((uint64_t)2 << i) * (uint64_t)b[i + 1];
        const unsigned blocks = (unsigned)((total + 255) / 256);
        hipLaunchKernelGGL(mps_dense_step_kernel, dim3(blocks), dim3(256), 0, hs, S, site_stride,
                           i, b[i], b[i + 1], cur, nxt, i == n - 1 ? f : 1.0);
        QS_HIP(hipGetLastError());
        cur = nxt;
    }
    return QS_OK;
}

int qs_mps_sample(const double* sites, const int32_t* bond, int64_t site_stride, int n,
                  int chi_max, const double* envs, const double* uniforms, int shots,
                  uint8_t* bits, void* stream) {
    clear_error();
    int st = check_chain(sites, bond, site_stride, n, chi_max);
    if (st) return st;
    QS_REQUIRE(shots >= 0, QS_E_ARG, "negative shot count");
    if (shots == 0) return QS_OK;
    QS_REQUIRE(envs != nullptr && uniforms != nullptr && bits != nullptr, QS_E_ARG,
               "null environments, uniforms or output");
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    const int blocks = (shots + kShotsPerBlock - 1) / kShotsPerBlock;
    hipLaunchKernelGGL(mps_sample_kernel, dim3(blocks), dim3(kShotLanes * kShotsPerBlock), 0, hs,
                       sites, bond, site_stride, n, chi_max,
                       reinterpret_cast<const cd2*>(envs), uniforms, shots, bits);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

}  // extern "C"
