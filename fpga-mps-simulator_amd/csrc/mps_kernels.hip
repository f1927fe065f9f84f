// MPS two-site update on gfx950: contraction (FP64 or complex FP32 on MFMA), one-sided
// Jacobi SVD in FP64, truncation, renormalisation and sqrt(S) absorption.
//
// Reference semantics: mps_core.MPSState._apply_adjacent_2q + _svd_truncate
// (mps_core.py:277-350) and the driver's copy of the same rule
// (fpga_mps_simulator.py:597-634); device contraction layout of
// contract_and_apply_2q (fpga_mps_simulator.cpp:287-371).
//
// One update = one workgroup per kernel:
//   contraction: theta[(a,s0),(s1,d)] = sum_pq G[2 s0 + s1, 2 p + q] sum_b A[a,p,b] B[b,q,d]
//                written as a (2chi_l x 2chi_r) complex128 matrix into the workspace;
//   svd:         up to chi 64: 64 teams of TL lanes (the autotuned team width); team t
//                holds two columns of theta in registers and orthogonalises pairs of
//                columns with complex Jacobi rotations in the odd-even ordering (even
//                steps inside a team, odd steps hand one column each way to the
//                neighbouring team through LDS); after convergence column norms are the
//                singular values, the kept columns (normalised) are U, and
//                Vh = S^-1 U^H theta is formed from theta in the workspace.  chi 128:
//                a blocked variant over a working copy of theta (mps_svd_big_kernel).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <mutex>
#include <utility>
#include <vector>

#include "qsim_internal.h"

namespace qs {

namespace {

constexpr int kMaxChi = 128;             // largest chi_max of a task
constexpr int kRegChi = 64;              // register SVD / MFMA contraction: 2*chi <= 128
constexpr int kMaxSweeps = 60;
// Stopping rule of the SVD kernels: a sweep in which every rotation had
// |a^H b| <= kFinalSweepRot ||a|| ||b|| is the last one.  Cyclic Jacobi converges
// quadratically at the end (on the chi-64 thetas the largest relative |a^H b| per sweep
// runs ... 1e-3, 1e-6, 1e-12, then the 3e-15 rounding floor), so what is left after such a
// sweep is far below the threshold at which rotations are applied at all, and the sweep
// that would only confirm it is not run.
constexpr double kFinalSweepRot = 1e-9;

__device__ __forceinline__ cd2 cmul(cd2 a, cd2 b) {
    return mk2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ cd2 cmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}
// acc += conj(a) * b
__device__ __forceinline__ cd2 cjmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
    return acc;
}

struct Dims {
    int cl, cm, cr;
};

// Device copy of a task: the caller's record plus its index in the caller's order
// (tasks are grouped by size class before upload).
struct DevTask {
    qs_mps_task t;
    int32_t orig;
    int32_t pad[3];
};

// Workspace elements per task: theta (2chi x 2chi); the chi-128 class adds its working
// copy of the columns.
__host__ __device__ inline size_t theta_slot_elems(int chi) {
    return (size_t)(chi > kRegChi ? 8 : 4) * chi * chi;
}

// Bond dimensions of the task's two sites, checked against its chi_max and against the
// workspace slot the launch gives each task (a task whose chi_max needs a larger slot
// is refused rather than written past its slot).
__device__ __forceinline__ bool load_dims(const qs_mps_task& T, Dims& d, size_t slot) {
    d.cl = T.bond[T.site];
    d.cm = T.bond[T.site + 1];
    d.cr = T.bond[T.site + 2];
    return d.cl >= 1 && d.cm >= 1 && d.cr >= 1 && d.cl <= T.chi_max && d.cm <= T.chi_max &&
           d.cr <= T.chi_max && T.chi_max >= 1 && T.chi_max <= kMaxChi &&
           theta_slot_elems(T.chi_max) <= slot;
}

__device__ __forceinline__ cd2* site_ptr(const qs_mps_task& T, int k) {
    return reinterpret_cast<cd2*>(T.sites) + (size_t)k * (size_t)T.site_stride;
}

// ------------------------------------------------------------------ FP64 contraction
__global__ void __launch_bounds__(256)
mps_contract_fp64_kernel(const DevTask* __restrict__ tasks, cd2* __restrict__ ws,
                         size_t theta_slot, int32_t* __restrict__ status) {
    const qs_mps_task& T = tasks[blockIdx.x].t;
    Dims d;
    if (!load_dims(T, d, theta_slot)) {
        if (threadIdx.x == 0 && status) status[tasks[blockIdx.x].orig] = QS_E_CHI;
        return;
    }
    const cd2* A = site_ptr(T, T.site);
    const cd2* B = site_ptr(T, T.site + 1);
    cd2* th = ws + (size_t)blockIdx.x * theta_slot;
    cd2 g[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = mk2(T.gate[2 * k], T.gate[2 * k + 1]);
    const int cl = d.cl, cm = d.cm, cr = d.cr;
    const int ldt = 2 * cr;
    for (int pr = threadIdx.x; pr < cl * cr; pr += blockDim.x) {
        const int a = pr / cr, e = pr - a * cr;
        cd2 t00 = mk2(0, 0), t01 = t00, t10 = t00, t11 = t00;
        const cd2* a0 = A + (size_t)(2 * a) * cm;
        const cd2* a1 = a0 + cm;
        for (int b = 0; b < cm; ++b) {
            const cd2 x0 = a0[b], x1 = a1[b];
            const cd2 y0 = B[(size_t)(2 * b) * cr + e], y1 = B[(size_t)(2 * b + 1) * cr + e];
            t00 = cmadd(t00, x0, y0);
            t01 = cmadd(t01, x0, y1);
            t10 = cmadd(t10, x1, y0);
            t11 = cmadd(t11, x1, y1);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            cd2 v = cmul(g[4 * s + 0], t00);
            v = cmadd(v, g[4 * s + 1], t01);
            v = cmadd(v, g[4 * s + 2], t10);
            v = cmadd(v, g[4 * s + 3], t11);
            const int s0 = s >> 1, s1 = s & 1;
            th[(size_t)(2 * a + s0) * ldt + s1 * cr + e] = v;
        }
    }
}

// ------------------------------------------------- complex FP32 contraction on MFMA
// GEMM T = A (M x K) * B (K x N), M = 2 chi_l rows (a,p), K = chi_m, N = 2 chi_r
// columns (q,d), in complex FP32 with FP32 accumulation on v_mfma_f32_16x16x4f32
// (4 real products per complex product).  Operands are staged through LDS as split
// real/imaginary FP32 planes; T goes back through LDS and the 4x4 gate mixes the
// (p,q) pairs in FP64 when theta is written.
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaThreads = 256;                 // 4 waves, one 64x64 quadrant each
constexpr int kMaxM = 2 * kRegChi;                // 128
constexpr int kMaxK = kRegChi;                    // 64

__global__ void __launch_bounds__(kMfmaThreads)
mps_contract_f32_mfma_kernel(const DevTask* __restrict__ tasks, cd2* __restrict__ ws,
                             size_t theta_slot, int32_t* __restrict__ status) {
    // Operand planes: A as [M][K] (re, im), B as [K][N] (re, im); K padded to 64,
    // M and N padded to 128.  2 * 128*64*4 B * 2 = 128 KiB; T reuses the same space.
    __shared__ float lds[2 * kMaxM * kMaxK * 2];
    float* Are = lds;
    float* Aim = Are + kMaxM * kMaxK;
    float* Bre = Aim + kMaxM * kMaxK;
    float* Bim = Bre + kMaxK * kMaxM;

    const qs_mps_task& T = tasks[blockIdx.x].t;
    Dims d;
    if (!load_dims(T, d, theta_slot) || d.cl > kRegChi || d.cm > kRegChi || d.cr > kRegChi) {
        if (threadIdx.x == 0 && status) status[tasks[blockIdx.x].orig] = QS_E_CHI;
        return;
    }
    const int M = 2 * d.cl, K = d.cm, N = 2 * d.cr;
    const int Kp = (K + 3) & ~3;
    const cd2* A = site_ptr(T, T.site);
    const cd2* B = site_ptr(T, T.site + 1);
    const int tid = threadIdx.x;

    for (int idx = tid; idx < kMaxM * Kp; idx += kMfmaThreads) {
        const int r = idx / Kp, k = idx - r * Kp;
        cd2 v = mk2(0, 0);
        if (r < M && k < K) v = A[(size_t)r * K + k];
        Are[r * kMaxK + k] = (float)v.x;
        Aim[r * kMaxK + k] = (float)v.y;
    }
    for (int idx = tid; idx < Kp * kMaxM; idx += kMfmaThreads) {
        const int k = idx / kMaxM, c = idx - k * kMaxM;
        cd2 v = mk2(0, 0);
        if (k < K && c < N) v = B[(size_t)k * N + c];
        Bre[k * kMaxM + c] = (float)v.x;
        Bim[k * kMaxM + c] = (float)v.y;
    }
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    const int qm = (wave >> 1) * 64, qn = (wave & 1) * 64;   // quadrant origin
    const int lr = lane & 15, lk = lane >> 4;                 // operand lane mapping
    floatx4 cre[4][4], cim[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cre[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
            cim[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    const bool active = qm < M && qn < N;
    if (active) {
        for (int k0 = 0; k0 < Kp; k0 += 4) {
            float ar[4], ai[4], nai[4], br[4], bi[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = qm + 16 * i + lr;
                ar[i] = Are[row * kMaxK + k0 + lk];
                ai[i] = Aim[row * kMaxK + k0 + lk];
                nai[i] = -ai[i];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = qn + 16 * j + lr;
                br[j] = Bre[(k0 + lk) * kMaxM + col];
                bi[j] = Bim[(k0 + lk) * kMaxM + col];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], br[j], cre[i][j], 0, 0, 0);
                    cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai[i], bi[j], cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bi[j], cim[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], br[j], cim[i][j], 0, 0, 0);
                }
        }
    }
    __syncthreads();
    // T back to LDS as [M][N] (re, im) planes: 2 * 128*128*4 B = 128 KiB.
    float* Tre = lds;
    float* Tim = lds + kMaxM * kMaxM;
    if (active) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    // 16x16x4 f32 output: lane holds rows 4*(lane/16)+e, column lane%16
                    const int row = qm + 16 * i + 4 * lk + e;
                    const int col = qn + 16 * j + lr;
                    Tre[row * kMaxM + col] = cre[i][j][e];
                    Tim[row * kMaxM + col] = cim[i][j][e];
                }
    }
    __syncthreads();
    cd2 g[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = mk2(T.gate[2 * k], T.gate[2 * k + 1]);
    cd2* th = ws + (size_t)blockIdx.x * theta_slot;
    const int cl = d.cl, cr = d.cr, ldt = 2 * cr;
    for (int pr = tid; pr < cl * cr; pr += kMfmaThreads) {
        const int a = pr / cr, e = pr - a * cr;
        const int r0 = 2 * a, r1 = 2 * a + 1, c0 = e, c1 = cr + e;
        const cd2 t00 = mk2(Tre[r0 * kMaxM + c0], Tim[r0 * kMaxM + c0]);
        const cd2 t01 = mk2(Tre[r0 * kMaxM + c1], Tim[r0 * kMaxM + c1]);
        const cd2 t10 = mk2(Tre[r1 * kMaxM + c0], Tim[r1 * kMaxM + c0]);
        const cd2 t11 = mk2(Tre[r1 * kMaxM + c1], Tim[r1 * kMaxM + c1]);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            cd2 v = cmul(g[4 * s + 0], t00);
            v = cmadd(v, g[4 * s + 1], t01);
            v = cmadd(v, g[4 * s + 2], t10);
            v = cmadd(v, g[4 * s + 3], t11);
            th[(size_t)(2 * a + (s >> 1)) * ldt + (s & 1) * cr + e] = v;
        }
    }
}

// The same complex-FP32 MFMA contraction for the chi 65..128 class, where theta (up to
// 256 x 256) is larger than one workgroup's LDS: workgroup (task, tile) owns the rows of
// a in [64 ta, 64 ta + 64) (both p) and the columns of d in [64 td, 64 td + 64) (both q),
// a 128 x 128 block of T, so the gate mixing of each (a, d) finds its four (p, q) entries
// in the same workgroup.  K = chi_m (up to 128) runs in chunks of 64 staged through LDS
// as in the kernel above; the accumulators stay in registers across chunks.
__global__ void __launch_bounds__(kMfmaThreads)
mps_contract_f32_mfma_big_kernel(const DevTask* __restrict__ tasks, cd2* __restrict__ ws,
                                 size_t theta_slot, int32_t* __restrict__ status) {
    __shared__ float lds[2 * kMaxM * kMaxK * 2];
    float* Are = lds;                           // [128 local rows][64 k]
    float* Aim = Are + kMaxM * kMaxK;
    float* Bre = Aim + kMaxM * kMaxK;           // [64 k][128 local columns]
    float* Bim = Bre + kMaxK * kMaxM;

    const qs_mps_task& T = tasks[blockIdx.x].t;
    Dims d;
    if (!load_dims(T, d, theta_slot)) {
        if (blockIdx.y == 0 && threadIdx.x == 0 && status)
            status[tasks[blockIdx.x].orig] = QS_E_CHI;
        return;
    }
    const int a0 = 64 * (blockIdx.y >> 1), d0 = 64 * (blockIdx.y & 1);
    if (a0 >= d.cl || d0 >= d.cr) return;        // uniform per workgroup
    const int K = d.cm, N = 2 * d.cr;
    const cd2* A = site_ptr(T, T.site);          // rows 2a + p, K columns
    const cd2* B = site_ptr(T, T.site + 1);      // K rows, columns q cr + d
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int qm = (wave >> 1) * 64, qn = (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    floatx4 cre[4][4], cim[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cre[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
            cim[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    for (int kc = 0; kc < K; kc += kMaxK) {
        const int kn = min(kMaxK, K - kc);
        const int Kp = (kn + 3) & ~3;
        for (int idx = tid; idx < kMaxM * Kp; idx += kMfmaThreads) {
            const int r = idx / Kp, k = idx - r * Kp;
            const int a = a0 + (r >> 1);
            cd2 v = mk2(0, 0);
            if (a < d.cl && k < kn) v = A[(size_t)(2 * a0 + r) * K + kc + k];
            Are[r * kMaxK + k] = (float)v.x;
            Aim[r * kMaxK + k] = (float)v.y;
        }
        for (int idx = tid; idx < Kp * kMaxM; idx += kMfmaThreads) {
            const int k = idx / kMaxM, c = idx - k * kMaxM;
            const int dd = d0 + (c & 63);
            cd2 v = mk2(0, 0);
            if (k < kn && dd < d.cr) v = B[(size_t)(kc + k) * N + (c >> 6) * d.cr + dd];
            Bre[k * kMaxM + c] = (float)v.x;
            Bim[k * kMaxM + c] = (float)v.y;
        }
        __syncthreads();
        for (int k0 = 0; k0 < Kp; k0 += 4) {
            float ar[4], ai[4], nai[4], br[4], bi[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = qm + 16 * i + lr;
                ar[i] = Are[row * kMaxK + k0 + lk];
                ai[i] = Aim[row * kMaxK + k0 + lk];
                nai[i] = -ai[i];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = qn + 16 * j + lr;
                br[j] = Bre[(k0 + lk) * kMaxM + col];
                bi[j] = Bim[(k0 + lk) * kMaxM + col];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], br[j], cre[i][j], 0, 0, 0);
                    cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai[i], bi[j], cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bi[j], cim[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], br[j], cim[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    float* Tre = lds;                             // [128 local rows][128 local columns]
    float* Tim = lds + kMaxM * kMaxM;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = qm + 16 * i + 4 * lk + e;
                const int col = qn + 16 * j + lr;
                Tre[row * kMaxM + col] = cre[i][j][e];
                Tim[row * kMaxM + col] = cim[i][j][e];
            }
    __syncthreads();
    cd2 g[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k] = mk2(T.gate[2 * k], T.gate[2 * k + 1]);
    cd2* th = ws + (size_t)blockIdx.x * theta_slot;
    const int cr = d.cr, ldt = 2 * cr;
    const int na = min(64, d.cl - a0), nd = min(64, cr - d0);
    for (int pr = tid; pr < na * nd; pr += kMfmaThreads) {
        const int al = pr / nd, dl = pr - al * nd;
        const int r0 = 2 * al, r1 = r0 + 1, c0 = dl, c1 = 64 + dl;
        const cd2 t00 = mk2(Tre[r0 * kMaxM + c0], Tim[r0 * kMaxM + c0]);
        const cd2 t01 = mk2(Tre[r0 * kMaxM + c1], Tim[r0 * kMaxM + c1]);
        const cd2 t10 = mk2(Tre[r1 * kMaxM + c0], Tim[r1 * kMaxM + c0]);
        const cd2 t11 = mk2(Tre[r1 * kMaxM + c1], Tim[r1 * kMaxM + c1]);
        const int a = a0 + al, e = d0 + dl;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            cd2 v = cmul(g[4 * s + 0], t00);
            v = cmadd(v, g[4 * s + 1], t01);
            v = cmadd(v, g[4 * s + 2], t10);
            v = cmadd(v, g[4 * s + 3], t11);
            th[(size_t)(2 * a + (s >> 1)) * ldt + (s & 1) * cr + e] = v;
        }
    }
}

// ------------------------------------------------------------------ Jacobi SVD
// FP64 reciprocal and reciprocal square root: the hardware estimate refined by two
// Newton steps (relative error at rounding level).
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}
__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double h = 0.5 * x * y;
        const double r = fma(-h, y, 0.5);
        y = fma(y, r, y);
    }
    return y;
}

// Complex Jacobi rotation zeroing a^H b for ||a||^2 = al, ||b||^2 = be, a^H b = gr + i gi:
//   a' = c a - s e^{-i phi} b,  b' = s e^{i phi} a + c b,  e^{i phi} = (gr + i gi)/|g|,
// with t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (be - al) / (2|g|),
// c = 1/sqrt(1 + t^2), s = c t.  Returns false (no rotation) when either column is a
// numerical zero or |g| <= tol ||a|| ||b||.  c^2 + s^2 and |e^{i phi}| equal 1 to rounding.
// tg = t |g| is what the rotation moves between the squared norms:
// ||a'||^2 = al - tg, ||b'||^2 = be + tg.
__device__ __forceinline__ bool jacobi_rotation(double al, double be, double gr, double gi,
                                                double zero_sq, double tol, double& c,
                                                double& s, double& er, double& ei,
                                                double& tg) {
    const double g2 = fma(gr, gr, gi * gi);
    if (!(al > zero_sq && be > zero_sq && g2 > tol * tol * al * be)) return false;
    const double rg = rsq_nr(g2);                    // 1 / |g|
    const double zeta = 0.5 * (be - al) * rg;
    double t;
    if (fabs(zeta) > 1e150) {
        t = 0.5 * rcp_nr(zeta);
    } else {
        const double q = fma(zeta, zeta, 1.0);
        const double root = q * rsq_nr(q);
        t = rcp_nr(fabs(zeta) + root);
        if (zeta < 0) t = -t;
    }
    c = rsq_nr(fma(t, t, 1.0));
    s = c * t;
    er = gr * rg;
    ei = gi * rg;
    tg = t * (g2 * rg);
    return true;
}

// A double moved across lanes by a DPP pattern (two 32-bit moves, no LDS traffic).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over the TL lanes of a team (TL a power of two, at most a wavefront).  Inside a
// 16-lane row the partners come from DPP patterns on the VALU: lane ^ 1 and lane ^ 2
// (quad permutes), then 7 - i and 15 - i (half-row and row mirrors), which pair the
// quads and half-rows once each holds a uniform partial sum; wider teams add the
// rows through cross-lane shuffles.  Partners add the same two values, so every lane
// of a team ends with bitwise the same sum.
template <int TL>
__device__ __forceinline__ double team_sum(double v) {
    if constexpr (TL >= 2) v += dpp_f64<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    if constexpr (TL >= 4) v += dpp_f64<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    if constexpr (TL >= 8) v += dpp_f64<0x141>(v);   // row_half_mirror
    if constexpr (TL >= 16) v += dpp_f64<0x140>(v);  // row_mirror
#pragma unroll
    for (int o = 16; o < TL; o *= 2) v += __shfl_xor(v, o, TL);
    return v;
}

// Size class CHI (8, 16, 32, 64): theta is at most 2CHI x 2CHI; CHI teams of TL lanes
// (one team per column pair), 2CHI/TL rows per lane, an LDS buffer of CHI columns of
// 2CHI rows (column exchange, then the kept left singular vectors).  The team width TL is
// the kernel's tile shape: narrower
// teams mean fewer threads per task and cheaper reductions, wider ones less work per
// lane; the launcher takes it from a per-class table (default 16, set by the autotuner).
template <int CHI>
struct SvdShared {
    static constexpr int LD = 2 * CHI;
    // column exchange of the sweeps, the Householder vectors of the preconditioning
    // QR and the kept left singular vectors; the transpose of R reuses it for R's
    // upper triangle packed by columns (tbuf)
    union {
        cd2 buf[CHI * LD];
        cd2 tbuf[CHI * (LD + 1)];
    };
    double qr_tau[2];
    double sval[2 * CHI];
    int order[2 * CHI];
    int rank[2 * CHI];
    double fro_part[CHI];
    double nslot[CHI];     // squared norm of the column in exchange slot t
    double fro;
    int rotated;
    int keep;
    double scale;
};

// Householder QR of the nr x nc matrix X held as in mps_svd_kernel (team t: columns 2t,
// 2t+1 in a, b; lane l: rows l + TL k), in place: afterwards the registers hold R
// (upper triangular, rows below the diagonal zero).  Step j: the team owning column j
// forms v = x - beta e_j (rows >= j of column j, beta = -e^{i arg x_j} ||x||, so no
// cancellation), writes R's column j and v into an LDS buffer (two buffers alternate,
// so one barrier per step), and every team applies H = I - (2 / v^H v) v v^H to its
// columns after j.  Column norms and dot products are team sums in fixed order.
template <int CHI, int TL>
__device__ __forceinline__ void householder_qr(cd2 (&a)[2 * CHI / TL], cd2 (&b)[2 * CHI / TL],
                                               SvdShared<CHI>& sh, int nr, int nc, int team,
                                               int l16) {
    constexpr int RPL = 2 * CHI / TL;
    constexpr int LD = 2 * CHI;
    const int steps = nr < nc ? nr : nc;
    for (int j = 0; j < steps; ++j) {
        cd2* vbuf = sh.buf + (j & 1) * LD;
        const int owner = j >> 1;
        if (team == owner) {
            cd2 (&x)[RPL] = (j & 1) ? b : a;
            double s2 = 0.0, ar = 0.0, ai = 0.0;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const int r = l16 + TL * k;
                if (r > j) s2 = fma(x[k].x, x[k].x, fma(x[k].y, x[k].y, s2));
                if (r == j) {
                    ar = x[k].x;
                    ai = x[k].y;
                }
            }
            s2 = team_sum<TL>(s2);
            ar = team_sum<TL>(ar);
            ai = team_sum<TL>(ai);
            const double aa = fma(ar, ar, ai * ai);
            const double nrm = sqrt(s2 + aa);
            double br = 0.0, bi = 0.0, tau2 = 0.0;
            if (nrm > 0.0) {
                if (aa > 0.0) {
                    const double f = -nrm / sqrt(aa);
                    br = f * ar;
                    bi = f * ai;
                } else {
                    br = -nrm;
                }
                const double vr = ar - br, vi = ai - bi;
                tau2 = 2.0 / (s2 + fma(vr, vr, vi * vi));
            }
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const int r = l16 + TL * k;
                cd2 v = mk2(0.0, 0.0);
                if (r > j) v = x[k];
                if (r == j) v = mk2(ar - br, ai - bi);
                vbuf[r] = v;
                if (r == j) x[k] = mk2(br, bi);
                if (r > j) x[k] = mk2(0.0, 0.0);
            }
            if (l16 == 0) sh.qr_tau[j & 1] = tau2;
        }
        __syncthreads();
        const double tau2 = sh.qr_tau[j & 1];
        if (tau2 != 0.0 && 2 * team + 1 > j && 2 * team < nc) {
            cd2 v[RPL];
#pragma unroll
            for (int k = 0; k < RPL; ++k) v[k] = vbuf[l16 + TL * k];
            auto apply = [&](cd2 (&x)[RPL]) __attribute__((always_inline)) {
                double wr = 0.0, wi = 0.0;       // w = v^H x
#pragma unroll
                for (int k = 0; k < RPL; ++k) {
                    wr = fma(v[k].x, x[k].x, fma(v[k].y, x[k].y, wr));
                    wi = fma(v[k].x, x[k].y, fma(-v[k].y, x[k].x, wi));
                }
                wr = team_sum<TL>(wr) * tau2;
                wi = team_sum<TL>(wi) * tau2;
#pragma unroll
                for (int k = 0; k < RPL; ++k) {   // x -= (tau2 w) v
                    x[k].x -= fma(wr, v[k].x, -wi * v[k].y);
                    x[k].y -= fma(wr, v[k].y, wi * v[k].x);
                }
            };
            if (2 * team > j) apply(a);
            if (2 * team + 1 > j && 2 * team + 1 < nc) apply(b);
        }
    }
    __syncthreads();
}

// Replaces the registers holding R (k0 x nc, upper trapezoidal, columns as in
// householder_qr) by L = R^H (nc x k0): team t receives columns 2t, 2t+1 of L, the
// conjugated rows 2t, 2t+1 of R, lane l rows l + TL k.  R's nonzero part (rows <= c of
// column c, at most (2 CHI)(2 CHI + 1) / 2 entries) goes to LDS packed by columns, so one
// round with one barrier moves all of it.
template <int CHI, int TL>
__device__ __forceinline__ void transpose_conj(cd2 (&a)[2 * CHI / TL], cd2 (&b)[2 * CHI / TL],
                                               SvdShared<CHI>& sh, int nc, int k0, int team,
                                               int l16) {
    constexpr int RPL = 2 * CHI / TL;
    static_assert(sizeof(sh.tbuf) / sizeof(cd2) >= (2 * CHI) * (2 * CHI + 1) / 2,
                  "packed triangle must fit");
    auto offset = [&](int c) {
        return c < k0 ? c * (c + 1) / 2 : k0 * (k0 + 1) / 2 + (c - k0) * k0;
    };
    if (2 * team < nc) {
        const int c0 = 2 * team, c1 = 2 * team + 1;
        const int o0 = offset(c0), o1 = offset(c1);
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const int r = l16 + TL * k;
            if (r <= c0 && r < k0) sh.tbuf[o0 + r] = a[k];
            if (c1 < nc && r <= c1 && r < k0) sh.tbuf[o1 + r] = b[k];
        }
    }
    __syncthreads();
    const int c0 = 2 * team, c1 = 2 * team + 1;   // columns of L = rows of R
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
        const int r = l16 + TL * k;                 // row of L = column of R
        cd2 u = mk2(0.0, 0.0), v = mk2(0.0, 0.0);
        if (r < nc) {
            const int o = offset(r);
            if (c0 < k0 && c0 <= r) u = sh.tbuf[o + c0];
            if (c1 < k0 && c1 <= r) v = sh.tbuf[o + c1];
        }
        a[k] = mk2(u.x, -u.y);
        b[k] = mk2(v.x, -v.y);
    }
    __syncthreads();
}

template <int CHI, int TL>
__global__ void __launch_bounds__(TL * CHI)
mps_svd_kernel(const DevTask* __restrict__ tasks, const cd2* __restrict__ ws,
               size_t theta_slot, int32_t* __restrict__ status) {
    constexpr int NT = TL * CHI;              // threads
    constexpr int RPL = 2 * CHI / TL;         // rows per lane (TL lanes x RPL = 2 CHI rows)
    static_assert(RPL >= 1 && RPL * TL == 2 * CHI && TL >= 2 && TL <= 64, "bad team width");
    constexpr int LD = 2 * CHI;               // LDS column stride
    __shared__ SvdShared<CHI> sh;
    const DevTask& DT = tasks[blockIdx.x];
    const qs_mps_task& T = DT.t;
    Dims d;
    if (!load_dims(T, d, theta_slot) || d.cl > CHI || d.cr > CHI) {
        if (threadIdx.x == 0 && status) status[DT.orig] = QS_E_CHI;
        return;
    }
    const int m = 2 * d.cl, n = 2 * d.cr;
    const int k0 = m < n ? m : n;                // singular values
    const int tid = threadIdx.x;
    const int team = tid / TL, l16 = tid % TL;
    const cd2* th = ws + (size_t)blockIdx.x * theta_slot;
    cd2* slot0 = sh.buf;                     // exchange slot of team t: column t

    // Preconditioning: the one-sided Jacobi runs on L = R^H, where theta^H = Q R
    // (Householder QR of theta^H, n x m).  theta = L Q^H with Q^H having orthonormal
    // rows, so theta and L have the same singular values and the same left singular
    // vectors; the columns of the triangular L are much closer to orthogonal than
    // theta's, and the sweeps drop from about 14 to about 10 at chi 64.  Vh is formed
    // from theta itself afterwards, as without preconditioning.
    //
    // X = theta^H in registers: team t holds columns 2t, 2t+1 of X (conjugated rows of
    // theta), lane l rows l + TL k of them.
    cd2 ca[RPL], cb[RPL];
    {
        const bool xact = team < m / 2;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const int r = l16 + TL * k;
            const cd2 u = (xact && r < n) ? th[(size_t)(2 * team) * n + r] : mk2(0, 0);
            const cd2 v = (xact && r < n) ? th[(size_t)(2 * team + 1) * n + r] : mk2(0, 0);
            ca[k] = mk2(u.x, -u.y);
            cb[k] = mk2(v.x, -v.y);
        }
    }
    householder_qr<CHI, TL>(ca, cb, sh, n, m, team, l16);
    // L = R^H: team t takes columns 2t, 2t+1 of L, i.e. conjugated rows 2t, 2t+1 of R.
    transpose_conj<CHI, TL>(ca, cb, sh, m, k0, team, l16);
    const int P = k0 / 2;                        // column pairs of L = active teams
    const bool act = team < P;
    // Frobenius norm^2 of theta (that of L): columns below 1e-14 * ||theta||_F are
    // numerical zeros and are not rotated.
    {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < RPL; ++k)
            s += ca[k].x * ca[k].x + ca[k].y * ca[k].y + cb[k].x * cb[k].x + cb[k].y * cb[k].y;
        s = team_sum<TL>(s);
        if (l16 == 0) sh.fro_part[team] = s;
        __syncthreads();
        if (tid == 0) {
            double f = 0.0;
            for (int t = 0; t < CHI; ++t) f += sh.fro_part[t];
            sh.fro = f;
        }
        __syncthreads();
    }
    const double zero_sq = 1e-28 * sh.fro;
    const double tol = 2.220446049250313e-16 * sqrt((double)m);

    // Jacobi rotation of the pair (x, y), x at the lower position: a' = c a - s e^{-i phi} b,
    // b' = s e^{i phi} a + c b with a = x, b = y.  nx, ny are the columns' squared norms
    // as tracked through the sweep (exact at its start, then moved by t |g| per rotation,
    // as LAPACK's zgesvj does): only a^H b is summed here.  A sweep without rotations
    // therefore tests every pair against exact norms.
    auto rotate = [&](cd2 (&x)[RPL], cd2 (&y)[RPL], double& nx, double& ny)
                      __attribute__((always_inline)) {
        double gr = 0, gi = 0;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            gr = fma(x[k].x, y[k].x, fma(x[k].y, y[k].y, gr));
            gi = fma(x[k].x, y[k].y, fma(-x[k].y, y[k].x, gi));
        }
        gr = team_sum<TL>(gr);
        gi = team_sum<TL>(gi);
        double c, sn, er, ei, tg;
        const bool big = fma(gr, gr, gi * gi) > kFinalSweepRot * kFinalSweepRot * nx * ny;
        if (jacobi_rotation(nx, ny, gr, gi, zero_sq, tol, c, sn, er, ei, tg)) {
            nx = fmax(nx - tg, 0.0);
            ny += tg;
            const cd2 sem = mk2(sn * er, -sn * ei);
            const cd2 sep = mk2(sn * er, sn * ei);
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const cd2 u = x[k], v = y[k];
                const cd2 sy = cmul(sem, v), sx = cmul(sep, u);
                x[k] = mk2(c * u.x - sy.x, c * u.y - sy.y);
                y[k] = mk2(c * v.x + sx.x, c * v.y + sx.y);
            }
            if (l16 == 0 && big) sh.rotated = 1;
        }
    };
    // Odd-even ordering.  Columns sit at positions 0..n-1, team t holding 2t and 2t+1.
    // Step s rotates the pairs (i, i+1), i = s mod 2 (mod 2), and the two columns of a
    // pair then trade positions; after n steps every pair has met exactly once (and the
    // order is reversed).  Even steps pair a team's own two columns: no exchange and no
    // barrier, and the trade only swaps which register holds which position.  Odd steps
    // pair position 2t+1 of team t with 2t+2 of team t+1: team t+1 hands that column
    // over through LDS, team t rotates the pair and hands the other column back (two
    // barriers, one column each way).  lo / hi are the registers holding positions 2t /
    // 2t+1.  Every team (idle ones too; slot t < CHI) parks lo in its slot and takes it
    // back at the end, team 0 unchanged, so lo holds no registers while team t works on
    // hi and the column it received.
    auto odd_step = [&](cd2 (&lo)[RPL], cd2 (&hi)[RPL], double& nlo, double& nhi)
                        __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < RPL; ++k) slot0[team * LD + l16 + TL * k] = lo[k];
        if (l16 == 0) sh.nslot[team] = nlo;
        __syncthreads();
        if (act && team < P - 1) {
            cd2 x[RPL];
#pragma unroll
            for (int k = 0; k < RPL; ++k) x[k] = slot0[(team + 1) * LD + l16 + TL * k];
            double nx = sh.nslot[team + 1];
            rotate(hi, x, nhi, nx);
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                slot0[(team + 1) * LD + l16 + TL * k] = hi[k];
                hi[k] = x[k];
            }
            if (l16 == 0) sh.nslot[team + 1] = nhi;
            nhi = nx;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPL; ++k) lo[k] = slot0[team * LD + l16 + TL * k];
        nlo = sh.nslot[team];
    };

    int sweeps = 0;
    for (; sweeps < kMaxSweeps; ++sweeps) {
        if (tid == 0) sh.rotated = 0;
        // exact squared norms at the start of every sweep
        double na = 0.0, nb = 0.0;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            na = fma(ca[k].x, ca[k].x, fma(ca[k].y, ca[k].y, na));
            nb = fma(cb[k].x, cb[k].x, fma(cb[k].y, cb[k].y, nb));
        }
        na = team_sum<TL>(na);
        nb = team_sum<TL>(nb);
        __syncthreads();
        // n = 2P steps: (even, odd) twice per trip with ca, cb as (lo, hi), then (hi, lo)
        for (int s = 0; s + 4 <= 2 * P; s += 4) {
            if (act) rotate(ca, cb, na, nb);
            odd_step(cb, ca, nb, na);
            if (act) rotate(cb, ca, nb, na);
            odd_step(ca, cb, na, nb);
        }
        if (P & 1) {
            // two steps left; they end with positions 2t, 2t+1 in cb, ca (the norms are
            // recomputed by the next sweep)
            if (act) rotate(ca, cb, na, nb);
            odd_step(cb, ca, nb, na);
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const cd2 x = ca[k];
                ca[k] = cb[k];
                cb[k] = x;
            }
        }
        __syncthreads();
        const int any = sh.rotated;
        __syncthreads();
        if (!any) break;
    }

    // Final column norms.  Team t owns column slots 2t (a) and 2t+1 (b); columns are
    // ranked by norm, so their original indices are not needed.
    double na = 0, nb = 0;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
        na = fma(ca[k].x, ca[k].x, fma(ca[k].y, ca[k].y, na));
        nb = fma(cb[k].x, cb[k].x, fma(cb[k].y, cb[k].y, nb));
    }
    na = sqrt(team_sum<TL>(na));
    nb = sqrt(team_sum<TL>(nb));
    if (act && l16 == 0) {
        sh.sval[2 * team] = na;
        sh.sval[2 * team + 1] = nb;
    }
    __syncthreads();
    if (tid < k0) {
        const double v = sh.sval[tid];
        int r = 0;
        for (int j = 0; j < k0; ++j) {
            const double w = sh.sval[j];
            r += (w > v) || (w == v && j < tid);
        }
        sh.rank[tid] = r;
        sh.order[r] = tid;
    }
    __syncthreads();
    if (tid == 0) {
        const int kmin = m < n ? m : n;
        const double s0 = sh.sval[sh.order[0]];
        const double thr = T.cutoff * s0;
        int cnt = 0;
        double total = 0.0;
        for (int k = 0; k < kmin; ++k) {
            const double s = sh.sval[sh.order[k]];
            cnt += (s > thr);
            total += s * s;
        }
        int keep = cnt < 1 ? 1 : cnt;
        if (keep > T.chi_max) keep = T.chi_max;
        double dlog = 0.0;
        if (keep < kmin) {
            double dropped = 0.0;
            for (int k = keep; k < kmin; ++k) {
                const double s = sh.sval[sh.order[k]];
                dropped += s * s;
            }
            if (dropped > 0.0) dlog += 0.5 * log(1.0 - dropped / fmax(total, 1e-300));
        }
        double kept = 0.0;
        for (int k = 0; k < keep; ++k) {
            const double s = sh.sval[sh.order[k]];
            kept += s * s;
        }
        const double nrm = sqrt(kept);
        double scale = 1.0;
        if (nrm > 1e-15) {
            dlog += log(nrm);
            scale = 1.0 / nrm;
        }
        sh.keep = keep;
        sh.scale = scale;
        *T.dlog = dlog;
        T.bond[T.site + 1] = keep;
        if (status) status[DT.orig] = (sweeps >= kMaxSweeps) ? QS_E_NOCONV : sweeps + 1;
    }
    __syncthreads();
    const int keep = sh.keep;
    const double scale = sh.scale;
    // Kept columns, normalised (U), into LDS at their rank: U[rank][row].
    if (act) {
        const int ra = sh.rank[2 * team], rb = sh.rank[2 * team + 1];
        const double ia = na > 0 ? 1.0 / na : 0.0, ib = nb > 0 ? 1.0 / nb : 0.0;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const int r = l16 + TL * k;
            if (ra < keep) sh.buf[ra * LD + r] = mk2(ca[k].x * ia, ca[k].y * ia);
            if (rb < keep) sh.buf[rb * LD + r] = mk2(cb[k].x * ib, cb[k].y * ib);
        }
    }
    __syncthreads();
    // A' = U sqrt(S_t), (chi_l, 2, keep) = m x keep, row-major.
    cd2* Ao = site_ptr(T, T.site);
    for (int idx = tid; idx < m * keep; idx += NT) {
        const int r = idx / keep, k = idx - r * keep;
        const double f = sqrt(sh.sval[sh.order[k]] * scale);
        const cd2 u = sh.buf[k * LD + r];
        Ao[idx] = mk2(u.x * f, u.y * f);
    }
    // B' = sqrt(S_t) Vh = sqrt(S_t)/S * U^H theta, (keep, 2, chi_r) = keep x n.
    cd2* Bo = site_ptr(T, T.site + 1);
    {
        constexpr int G = NT / LD;                  // groups of rows of B' (TL / 2)
        constexpr int QN = CHI / G > 0 ? CHI / G : 1;   // rows per group (keep <= CHI)
        const int c = tid % LD, kg = tid / LD;
        if (c < n) {
            cd2 acc[QN];
#pragma unroll
            for (int q = 0; q < QN; ++q) acc[q] = mk2(0, 0);
            for (int r = 0; r < m; ++r) {
                const cd2 t = th[(size_t)r * n + c];
#pragma unroll
                for (int q = 0; q < QN; ++q) {
                    const int k = kg + G * q;
                    if (k < keep) acc[q] = cjmadd(acc[q], sh.buf[k * LD + r], t);
                }
            }
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                const int k = kg + G * q;
                if (k < keep) {
                    const double s = sh.sval[sh.order[k]];
                    const double f = s > 0 ? sqrt(s * scale) / s : 0.0;
                    Bo[(size_t)k * n + c] = mk2(acc[q].x * f, acc[q].y * f);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ one-site gate
__global__ void mps_apply_1q_kernel(cd2* __restrict__ sites, const int32_t* __restrict__ bond,
                                    int64_t stride, int site, cd2 m00, cd2 m01,
                                    cd2 m10, cd2 m11) {
    const int cl = bond[site], cr = bond[site + 1];
    cd2* A = sites + (size_t)site * stride;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < cl * cr;
         idx += gridDim.x * blockDim.x) {
        const int a = idx / cr, b = idx - a * cr;
        cd2* p0 = A + (size_t)(2 * a) * cr + b;
        cd2* p1 = p0 + cr;
        const cd2 x0 = *p0, x1 = *p1;
        *p0 = cmadd(cmul(m00, x0), m01, x1);
        *p1 = cmadd(cmul(m10, x0), m11, x1);
    }
}

// Many one-site gates, one workgroup per gate; the gates of one launch act on distinct
// sites (the caller composes gates that share a site), so workgroups never overlap.
struct OneSiteGate {
    int32_t site, pad[3];
    double m[8];
};
constexpr int kMax1qPerLaunch = 32;
struct OneSiteBatch {
    OneSiteGate g[kMax1qPerLaunch];
};

__global__ void __launch_bounds__(256)
mps_apply_1q_batch_kernel(cd2* __restrict__ sites, const int32_t* __restrict__ bond,
                          int64_t stride, OneSiteBatch batch) {
    const OneSiteGate& G = batch.g[blockIdx.x];
    const int site = G.site;
    const int cl = bond[site], cr = bond[site + 1];
    const cd2 m00 = mk2(G.m[0], G.m[1]), m01 = mk2(G.m[2], G.m[3]);
    const cd2 m10 = mk2(G.m[4], G.m[5]), m11 = mk2(G.m[6], G.m[7]);
    cd2* A = sites + (size_t)site * stride;
    for (int idx = threadIdx.x; idx < cl * cr; idx += blockDim.x) {
        const int a = idx / cr, b = idx - a * cr;
        cd2* p0 = A + (size_t)(2 * a) * cr + b;
        cd2* p1 = p0 + cr;
        const cd2 x0 = *p0, x1 = *p1;
        *p0 = cmadd(cmul(m00, x0), m01, x1);
        *p1 = cmadd(cmul(m10, x0), m11, x1);
    }
}

// Reset of one site (fpga_mps_simulator.py:530-552): zero the sigma = 1 half, divide the
// site by its norm when that exceeds 1e-15, otherwise make it the |0> tensor t[0,0,0] = 1.
// One workgroup; the squared norm is summed in index order per thread and folded by a
// fixed tree.
__global__ void __launch_bounds__(256)
mps_reset_site_kernel(cd2* __restrict__ sites, const int32_t* __restrict__ bond,
                      int64_t stride, int site) {
    __shared__ double red[256];
    const int cl = bond[site], cr = bond[site + 1];
    cd2* A = sites + (size_t)site * stride;
    const int t = threadIdx.x;
    double part = 0.0;
    for (int idx = t; idx < cl * cr; idx += 256) {
        const int a = idx / cr, b = idx - a * cr;
        const cd2 x = A[(size_t)(2 * a) * cr + b];
        part = fma(x.x, x.x, fma(x.y, x.y, part));
        A[(size_t)(2 * a + 1) * cr + b] = mk2(0.0, 0.0);
    }
    red[t] = part;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    const double nrm = sqrt(red[0]);
    const bool keep = nrm > 1e-15;
    const double inv = keep ? 1.0 / nrm : 0.0;
    for (int idx = t; idx < cl * cr; idx += 256) {
        const int a = idx / cr, b = idx - a * cr;
        cd2* p = A + (size_t)(2 * a) * cr + b;
        const cd2 x = *p;
        *p = keep ? mk2(x.x * inv, x.y * inv) : mk2(idx == 0 ? 1.0 : 0.0, 0.0);
    }
}

}  // namespace
}  // namespace qs

using namespace qs;

// ------------------------------------------------- Jacobi SVD, chi_max up to 128
// theta up to 256 x 256 (1 MiB) does not fit on chip, so the columns live in a working
// copy W (column-major, next to theta in the task's workspace slot, L2-resident while
// the task runs) and travel through LDS in blocks of kBlkCols columns (below).  Pairings
// follow the round-robin tournament: in round r the player in slot s >= 1 is
// 1 + (s - 1 + r) mod (2P - 1), slot 0 keeps player 0, and pair i is
// (slot i, slot 2P - 1 - i), so every player meets every other once per 2P - 1 rounds.
// 1024 threads; the Frobenius norm and the final column norms use 32 teams of 32 lanes.
constexpr int kBigLanes = 32;
constexpr int kBigTeams = 32;
constexpr int kBigThreads = kBigTeams * kBigLanes;    // 1024

__device__ __forceinline__ double big_sum(double v) { return team_sum<kBigLanes>(v); }

// Column blocks of the chi-128 class: 32 columns of up to 256 rows.  In a block pair the
// first block sits in registers, one column per 32-lane half-wave (8 rows per lane), the
// second in LDS (column stride padded by two entries, so the two half-waves of a wave
// read different banks); inside a block (the tournament's first round) pairs are rotated
// by whole waves on the block in LDS (4 rows per lane).
constexpr int kBlkCols = 32;
constexpr int kBlkRows = 2 * kMaxChi;              // 256
constexpr int kBlkLd = kBlkRows + 2;
constexpr int kHalfLanes = 32;                     // lanes per register column
constexpr int kHalfRpl = kBlkRows / kHalfLanes;    // 8
constexpr int kWaveRpl = kBlkRows / 64;            // 4

__device__ __forceinline__ int big_player(int slot, int r, int P) {
    return slot == 0 ? 0 : 1 + (slot - 1 + r) % (2 * P - 1);
}

struct BigShared {
    double bn[kBlkCols];       // squared norms of the LDS block's columns
    double sval[2 * kMaxChi];
    int order[2 * kMaxChi];
    int rank[2 * kMaxChi];
    double part[kBigTeams];
    double fro;
    int rotated;
    int keep;
    double scale;
};

// One-sided Jacobi rotation of the column pair (x, y) held by a team of TL lanes (RPL rows
// per lane): sums a^H b, rotates when |a^H b| > tol ||a|| ||b|| (jacobi_rotation) and
// moves t |g| between the tracked squared norms.  True if the pair was rotated; any_big is
// set when the rotation was above the stopping threshold.
template <int TL, int RPL>
__device__ __forceinline__ bool rotate_team(cd2 (&x)[RPL], cd2 (&y)[RPL], double& nx,
                                            double& ny, double zero_sq, double tol,
                                            bool& any_big) {
    double gr = 0, gi = 0, gr2 = 0, gi2 = 0;
#pragma unroll
    for (int k = 0; k < RPL; k += 2) {
        gr = fma(x[k].x, y[k].x, fma(x[k].y, y[k].y, gr));
        gi = fma(x[k].x, y[k].y, fma(-x[k].y, y[k].x, gi));
        if (k + 1 < RPL) {
            gr2 = fma(x[k + 1].x, y[k + 1].x, fma(x[k + 1].y, y[k + 1].y, gr2));
            gi2 = fma(x[k + 1].x, y[k + 1].y, fma(-x[k + 1].y, y[k + 1].x, gi2));
        }
    }
    gr = team_sum<TL>(gr + gr2);
    gi = team_sum<TL>(gi + gi2);
    double c, sn, er, ei, tg;
    const bool big = fma(gr, gr, gi * gi) > kFinalSweepRot * kFinalSweepRot * nx * ny;
    if (!jacobi_rotation(nx, ny, gr, gi, zero_sq, tol, c, sn, er, ei, tg)) return false;
    if (big) any_big = true;
    nx = fmax(nx - tg, 0.0);
    ny += tg;
    const cd2 sem = mk2(sn * er, -sn * ei);
    const cd2 sep = mk2(sn * er, sn * ei);
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
        const cd2 u = x[k], v = y[k];
        const cd2 sy = cmul(sem, v), sx = cmul(sep, u);
        x[k] = mk2(c * u.x - sy.x, c * u.y - sy.y);
        y[k] = mk2(c * v.x + sx.x, c * v.y + sx.y);
    }
    return true;
}

// Householder QR of the n x n matrix X in W (column-major, leading dimension n) in panels
// of kBlkCols columns, then W := L = R^H in place.  Panel p sits in registers (half-wave
// h: column 32 p + h, 8 rows per lane); its 32 reflections are formed one after the other
// (half-wave j forms v = x - beta e_jj from its column, beta = -e^{i arg x_jj} ||x||,
// writes v into row j of V in LDS and R's column into its registers; the later columns
// of the panel apply H_j = I - tau_j v v^H at once, one barrier per reflection), then
// every trailing block of 32 columns is loaded, takes the panel's 32 reflections in
// order and is written back: W is read and written once per panel and trailing block.
// tau lives in `taus` (kBlkCols doubles of LDS), V in `vbuf` (kBlkCols x kBlkLd).
__device__ __noinline__ void big_lq_precondition(cd2* __restrict__ W, int n,
                                                 cd2* __restrict__ vbuf,
                                                 double* __restrict__ taus, int tid) {
    const int half = tid / kHalfLanes, hl = tid % kHalfLanes;
    const int nbk = (n + kBlkCols - 1) / kBlkCols;
    auto reflect = [&](cd2 (&x)[kHalfRpl], int j) __attribute__((always_inline)) {
        const double tau2 = taus[j];
        if (tau2 == 0.0) return;
        const cd2* v = vbuf + j * kBlkLd;
        double wr = 0.0, wi = 0.0;
#pragma unroll
        for (int k = 0; k < kHalfRpl; ++k) {
            const cd2 vk = v[hl + kHalfLanes * k];
            wr = fma(vk.x, x[k].x, fma(vk.y, x[k].y, wr));
            wi = fma(vk.x, x[k].y, fma(-vk.y, x[k].x, wi));
        }
        wr = team_sum<kHalfLanes>(wr) * tau2;
        wi = team_sum<kHalfLanes>(wi) * tau2;
#pragma unroll
        for (int k = 0; k < kHalfRpl; ++k) {
            const cd2 vk = v[hl + kHalfLanes * k];
            x[k].x -= fma(wr, vk.x, -wi * vk.y);
            x[k].y -= fma(wr, vk.y, wi * vk.x);
        }
    };
    for (int p = 0; p < nbk; ++p) {
        const int col = p * kBlkCols + half;
        cd2 x[kHalfRpl];
#pragma unroll
        for (int k = 0; k < kHalfRpl; ++k) {
            const int r = hl + kHalfLanes * k;
            x[k] = (col < n && r < n) ? W[(size_t)col * n + r] : mk2(0, 0);
        }
        for (int j = 0; j < kBlkCols; ++j) {
            const int jj = p * kBlkCols + j;
            if (jj >= n) break;
            if (half == j) {
                double s2 = 0.0, ar = 0.0, ai = 0.0;
#pragma unroll
                for (int k = 0; k < kHalfRpl; ++k) {
                    const int r = hl + kHalfLanes * k;
                    if (r > jj) s2 = fma(x[k].x, x[k].x, fma(x[k].y, x[k].y, s2));
                    if (r == jj) {
                        ar = x[k].x;
                        ai = x[k].y;
                    }
                }
                s2 = team_sum<kHalfLanes>(s2);
                ar = team_sum<kHalfLanes>(ar);
                ai = team_sum<kHalfLanes>(ai);
                const double aa = fma(ar, ar, ai * ai);
                const double nrm = sqrt(s2 + aa);
                double br = 0.0, bi = 0.0, tau2 = 0.0;
                if (nrm > 0.0) {
                    if (aa > 0.0) {
                        const double f = -nrm / sqrt(aa);
                        br = f * ar;
                        bi = f * ai;
                    } else {
                        br = -nrm;
                    }
                    const double vr = ar - br, vi = ai - bi;
                    tau2 = 2.0 / (s2 + fma(vr, vr, vi * vi));
                }
#pragma unroll
                for (int k = 0; k < kHalfRpl; ++k) {
                    const int r = hl + kHalfLanes * k;
                    cd2 v = mk2(0.0, 0.0);
                    if (r > jj) v = x[k];
                    if (r == jj) v = mk2(ar - br, ai - bi);
                    vbuf[j * kBlkLd + r] = v;
                    if (r == jj) x[k] = mk2(br, bi);
                    if (r > jj) x[k] = mk2(0.0, 0.0);
                }
                if (hl == 0) taus[j] = tau2;
            }
            __syncthreads();
            if (half > j && col < n) reflect(x, j);
        }
#pragma unroll
        for (int k = 0; k < kHalfRpl; ++k) {
            const int r = hl + kHalfLanes * k;
            if (col < n && r < n) W[(size_t)col * n + r] = x[k];
        }
        const int nref = n - p * kBlkCols < kBlkCols ? n - p * kBlkCols : kBlkCols;
        for (int q = p + 1; q < nbk; ++q) {
            const int cq = q * kBlkCols + half;
            if (cq >= n) continue;
            cd2 y[kHalfRpl];
#pragma unroll
            for (int k = 0; k < kHalfRpl; ++k) {
                const int r = hl + kHalfLanes * k;
                y[k] = r < n ? W[(size_t)cq * n + r] : mk2(0, 0);
            }
            for (int j = 0; j < nref; ++j) reflect(y, j);
#pragma unroll
            for (int k = 0; k < kHalfRpl; ++k) {
                const int r = hl + kHalfLanes * k;
                if (r < n) W[(size_t)cq * n + r] = y[k];
            }
        }
        __syncthreads();   // V and tau are rewritten by the next panel
    }
    // W := R^H in place (R upper triangular, zero below the diagonal), by 32 x 32 tile
    // pairs through LDS: tiles (I, J) and (J, I) are read, conjugated and written back
    // transposed.  Thread (i, j) of the 1024 handles entry (i, j) of both tiles.
    cd2* ta = vbuf;
    cd2* tb = vbuf + kBlkCols * (kBlkCols + 1);
    const int ti = tid / kBlkCols, tj = tid % kBlkCols;
    for (int I = 0; I < nbk; ++I) {
        for (int J = I; J < nbk; ++J) {
            // entry (row, col) of W = R[row][col]; tile (I, J) covers rows 32 I.., cols 32 J..
            const int r1 = I * kBlkCols + tj, c1 = J * kBlkCols + ti;
            const int r2 = J * kBlkCols + tj, c2 = I * kBlkCols + ti;
            ta[ti * (kBlkCols + 1) + tj] = (r1 < n && c1 < n) ? W[(size_t)c1 * n + r1] : mk2(0, 0);
            tb[ti * (kBlkCols + 1) + tj] = (r2 < n && c2 < n) ? W[(size_t)c2 * n + r2] : mk2(0, 0);
            __syncthreads();
            // L[r][c] = conj(R[c][r]): tile (I, J) of L comes from tile (J, I) of R
            const cd2 u = tb[tj * (kBlkCols + 1) + ti];   // R[32 J + ti][32 I + tj]
            const cd2 v = ta[tj * (kBlkCols + 1) + ti];   // R[32 I + ti][32 J + tj]
            if (r1 < n && c1 < n) W[(size_t)c1 * n + r1] = mk2(u.x, -u.y);   // L[r1][c1]
            if (r2 < n && c2 < n) W[(size_t)c2 * n + r2] = mk2(v.x, -v.y);   // L[r2][c2]
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(kBigThreads)
mps_svd_big_kernel(const DevTask* __restrict__ tasks, cd2* __restrict__ ws, size_t slot,
                   int32_t* __restrict__ status) {
    __shared__ BigShared sh;
    __shared__ cd2 blk[kBlkCols * kBlkLd];       // the LDS block of a block pair (129 KiB)
    const DevTask& DT = tasks[blockIdx.x];
    const qs_mps_task& T = DT.t;
    Dims d;
    if (!load_dims(T, d, slot)) {
        if (threadIdx.x == 0 && status) status[DT.orig] = QS_E_CHI;
        return;
    }
    const int m = 2 * d.cl, n = 2 * d.cr;
    const int tid = threadIdx.x;
    const int team = tid / kBigLanes, l16 = tid % kBigLanes;
    const cd2* th = ws + (size_t)blockIdx.x * slot;
    cd2* W = ws + (size_t)blockIdx.x * slot + slot / 2;

    // Square theta (m == n, every update at full bond) is preconditioned as in the
    // register kernel: W starts as X = theta^H (column c of X is the conjugated row c of
    // theta, so W is the conjugated theta entry by entry), is factored X = Q R in place
    // and replaced by L = R^H, on which the sweeps run (same singular values and left
    // singular vectors as theta, about half the sweeps at chi 128).  Other shapes run on
    // theta's columns directly.
    const bool precond = m == n;
    double fpart = 0.0;
    for (int idx = tid; idx < m * n; idx += kBigThreads) {
        const int r = idx / n, c = idx - r * n;
        const cd2 v = th[idx];
        if (precond)
            W[idx] = mk2(v.x, -v.y);
        else
            W[(size_t)c * m + r] = v;
        fpart = fma(v.x, v.x, fma(v.y, v.y, fpart));
    }
    fpart = big_sum(fpart);
    if (l16 == 0) sh.part[team] = fpart;
    __syncthreads();
    if (tid == 0) {
        double f = 0.0;
        for (int t = 0; t < kBigTeams; ++t) f += sh.part[t];
        sh.fro = f;
    }
    __syncthreads();
    const double zero_sq = 1e-28 * sh.fro;
    const double tol = 2.220446049250313e-16 * sqrt((double)m);
    if (precond) big_lq_precondition(W, n, blk, sh.bn, tid);

    // Blocked one-sided Jacobi.  theta's columns are cut into blocks of kBlkCols; a
    // sweep is a round-robin tournament over the blocks.  In the tournament's first
    // round every block appears once: its inner pairs are rotated on it in LDS (31
    // cyclic rounds, wave w takes pair w).  Then, for each block pair of every round,
    // the first block goes to registers (half-wave h holds its column h) and the second
    // to LDS, and the 32 x 32 cross pairs are rotated in 32 cyclic rounds (half-wave h
    // meets LDS column (h + k) mod 32 in round k); both blocks go back to W.  W is read
    // and written once per block pair: about 2 (nb - 1) passes over theta per sweep.
    const int nb = (n + kBlkCols - 1) / kBlkCols;
    const int nbp = nb + (nb & 1);                // even, a dummy block when nb is odd
    const int PB = nbp / 2;
    const int wave = tid / 64, lane = tid % 64;
    const int half = tid / kHalfLanes, hl = tid % kHalfLanes;   // 32 half-waves
    auto load_lds = [&](int blkid) {
        for (int idx = tid; idx < kBlkCols * kBlkRows; idx += kBigThreads) {
            const int i = idx / kBlkRows, r = idx - i * kBlkRows;
            const int col = blkid * kBlkCols + i;
            blk[i * kBlkLd + r] = (col < n && r < m) ? W[(size_t)col * m + r] : mk2(0, 0);
        }
    };
    auto store_lds = [&](int blkid) {
        for (int idx = tid; idx < kBlkCols * kBlkRows; idx += kBigThreads) {
            const int i = idx / kBlkRows, r = idx - i * kBlkRows;
            const int col = blkid * kBlkCols + i;
            if (col < n && r < m) W[(size_t)col * m + r] = blk[i * kBlkLd + r];
        }
    };
    // exact squared norms of the LDS block's columns (half-wave h: column h)
    auto lds_norms = [&]() {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < kHalfRpl; ++k) {
            const cd2 z = blk[half * kBlkLd + hl + kHalfLanes * k];
            v = fma(z.x, z.x, fma(z.y, z.y, v));
        }
        v = team_sum<kHalfLanes>(v);
        if (hl == 0) sh.bn[half] = v;
    };
    bool any_big = false;

    int sweeps = 0;
    for (; sweeps < kMaxSweeps; ++sweeps) {
        if (tid == 0) sh.rotated = 0;
        __syncthreads();
        any_big = false;
        for (int ro = 0; ro < nbp - 1; ++ro) {
            for (int pi = 0; pi < PB; ++pi) {
                const int bi = big_player(pi, ro, PB), bj = big_player(nbp - 1 - pi, ro, PB);
                if (ro == 0) {
                    // pairs inside each real block of the pair, on the block in LDS
                    for (int h2 = 0; h2 < 2; ++h2) {
                        const int b = h2 == 0 ? bi : bj;
                        if (b >= nb) continue;
                        load_lds(b);
                        __syncthreads();
                        lds_norms();
                        __syncthreads();
                        for (int rd = 0; rd < kBlkCols - 1; ++rd) {
                            const int ca = big_player(wave, rd, kBlkCols / 2);
                            const int cb = big_player(kBlkCols - 1 - wave, rd, kBlkCols / 2);
                            cd2 x[kWaveRpl], y[kWaveRpl];
#pragma unroll
                            for (int k = 0; k < kWaveRpl; ++k) {
                                x[k] = blk[ca * kBlkLd + lane + 64 * k];
                                y[k] = blk[cb * kBlkLd + lane + 64 * k];
                            }
                            double nx = sh.bn[ca], ny = sh.bn[cb];
                            if (rotate_team<64, kWaveRpl>(x, y, nx, ny, zero_sq, tol, any_big)) {
#pragma unroll
                                for (int k = 0; k < kWaveRpl; ++k) {
                                    blk[ca * kBlkLd + lane + 64 * k] = x[k];
                                    blk[cb * kBlkLd + lane + 64 * k] = y[k];
                                }
                                if (lane == 0) {
                                    sh.bn[ca] = nx;
                                    sh.bn[cb] = ny;
                                }
                            }
                            __syncthreads();
                        }
                        store_lds(b);
                        __syncthreads();
                    }
                }
                if (bi >= nb || bj >= nb) continue;
                // first block into registers (half-wave h: column h), second into LDS
                cd2 x[kHalfRpl];
                {
                    const int col = bi * kBlkCols + half;
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k < kHalfRpl; ++k) {
                        const int r = hl + kHalfLanes * k;
                        x[k] = (col < n && r < m) ? W[(size_t)col * m + r] : mk2(0, 0);
                        v = fma(x[k].x, x[k].x, fma(x[k].y, x[k].y, v));
                    }
                    double nx = team_sum<kHalfLanes>(v);
                    load_lds(bj);
                    __syncthreads();
                    lds_norms();
                    __syncthreads();
                    for (int kk = 0; kk < kBlkCols; ++kk) {
                        const int jc = (half + kk) % kBlkCols;
                        cd2 y[kHalfRpl];
#pragma unroll
                        for (int k = 0; k < kHalfRpl; ++k) y[k] = blk[jc * kBlkLd + hl + kHalfLanes * k];
                        double ny = sh.bn[jc];
                        if (rotate_team<kHalfLanes, kHalfRpl>(x, y, nx, ny, zero_sq, tol, any_big)) {
#pragma unroll
                            for (int k = 0; k < kHalfRpl; ++k)
                                blk[jc * kBlkLd + hl + kHalfLanes * k] = y[k];
                            if (hl == 0) sh.bn[jc] = ny;
                        }
                        __syncthreads();
                    }
#pragma unroll
                    for (int k = 0; k < kHalfRpl; ++k) {
                        const int r = hl + kHalfLanes * k;
                        if (col < n && r < m) W[(size_t)col * m + r] = x[k];
                    }
                }
                store_lds(bj);
                __syncthreads();
            }
        }
        if (any_big) sh.rotated = 1;
        __syncthreads();
        const int any = sh.rotated;
        __syncthreads();
        if (!any) break;
    }

    // Column norms (singular values), ranking, keep rule and norm_log as in the
    // register kernel.
    for (int c = team; c < n; c += kBigTeams) {
        double v = 0.0;
        for (int r = l16; r < m; r += kBigLanes) {
            const cd2 x = W[(size_t)c * m + r];
            v = fma(x.x, x.x, fma(x.y, x.y, v));
        }
        v = big_sum(v);
        if (l16 == 0) sh.sval[c] = sqrt(v);
    }
    __syncthreads();
    if (tid < n) {
        const double v = sh.sval[tid];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const double w = sh.sval[j];
            r += (w > v) || (w == v && j < tid);
        }
        sh.rank[tid] = r;
        sh.order[r] = tid;
    }
    __syncthreads();
    if (tid == 0) {
        const int kmin = m < n ? m : n;
        const double s0 = sh.sval[sh.order[0]];
        const double thr = T.cutoff * s0;
        int cnt = 0;
        double total = 0.0;
        for (int k = 0; k < kmin; ++k) {
            const double sv = sh.sval[sh.order[k]];
            cnt += (sv > thr);
            total += sv * sv;
        }
        int keep = cnt < 1 ? 1 : cnt;
        if (keep > T.chi_max) keep = T.chi_max;
        double dlog = 0.0;
        if (keep < kmin) {
            double dropped = 0.0;
            for (int k = keep; k < kmin; ++k) {
                const double sv = sh.sval[sh.order[k]];
                dropped += sv * sv;
            }
            if (dropped > 0.0) dlog += 0.5 * log(1.0 - dropped / fmax(total, 1e-300));
        }
        double kept = 0.0;
        for (int k = 0; k < keep; ++k) {
            const double sv = sh.sval[sh.order[k]];
            kept += sv * sv;
        }
        const double nrm = sqrt(kept);
        double scale = 1.0;
        if (nrm > 1e-15) {
            dlog += log(nrm);
            scale = 1.0 / nrm;
        }
        sh.keep = keep;
        sh.scale = scale;
        *T.dlog = dlog;
        T.bond[T.site + 1] = keep;
        if (status) status[DT.orig] = (sweeps >= kMaxSweeps) ? QS_E_NOCONV : sweeps + 1;
    }
    __syncthreads();
    const int keep = sh.keep;
    const double scale = sh.scale;
    // A' = U sqrt(S_t): column k of U is W[:, order[k]] / s_k.
    cd2* Ao = site_ptr(T, T.site);
    for (int idx = tid; idx < m * keep; idx += kBigThreads) {
        const int r = idx / keep, k = idx - r * keep;
        const double sv = sh.sval[sh.order[k]];
        const double f = sv > 0 ? sqrt(sv * scale) / sv : 0.0;
        const cd2 u = W[(size_t)sh.order[k] * m + r];
        Ao[idx] = mk2(u.x * f, u.y * f);
    }
    // B' = sqrt(S_t) / S * U^H theta = sqrt(S_t) / S^2 * W_k^H theta, in passes of
    // kBlkCols kept columns: the pass's columns of W go to LDS, thread (c, g) (column c of
    // theta, group g of 1024 / 256) accumulates 8 of them over the rows with theta[r][c]
    // read once per pass (theta is read about keep / 8 times instead of keep times).
    cd2* Bo = site_ptr(T, T.site + 1);
    constexpr int kGroups = kBigThreads / kBlkRows;            // 4
    constexpr int kPerGroup = kBlkCols / kGroups;              // 8
    for (int k0 = 0; k0 < keep; k0 += kBlkCols) {
        __syncthreads();
        for (int idx = tid; idx < kBlkCols * kBlkRows; idx += kBigThreads) {
            const int i = idx / kBlkRows, r = idx - i * kBlkRows;
            const int k = k0 + i;
            blk[i * kBlkLd + r] = (k < keep && r < m) ? W[(size_t)sh.order[k] * m + r] : mk2(0, 0);
        }
        __syncthreads();
        const int c = tid % kBlkRows, g = tid / kBlkRows;
        if (c < n) {
            cd2 acc[kPerGroup];
#pragma unroll
            for (int q = 0; q < kPerGroup; ++q) acc[q] = mk2(0, 0);
            for (int r = 0; r < m; ++r) {
                const cd2 t = th[(size_t)r * n + c];
#pragma unroll
                for (int q = 0; q < kPerGroup; ++q)
                    acc[q] = cjmadd(acc[q], blk[(g * kPerGroup + q) * kBlkLd + r], t);
            }
#pragma unroll
            for (int q = 0; q < kPerGroup; ++q) {
                const int k = k0 + g * kPerGroup + q;
                if (k < keep) {
                    const double sv = sh.sval[sh.order[k]];
                    const double f = sv > 0 ? sqrt(sv * scale) / (sv * sv) : 0.0;
                    Bo[(size_t)k * n + c] = mk2(acc[q].x * f, acc[q].y * f);
                }
            }
        }
    }
}

// Device scratch of qs_mps_update_batch_static: one batch's task table and theta slots
// (the workspace the other entry points take as an argument).  Each device the library
// runs on holds its own copy until the process ends.  In host code `scratch` is only a
// symbol: its device address on the current device comes from hipGetSymbolAddress and is
// kept per device in g_scratch_ptr.
static __device__ __align__(256) char scratch[QS_MPS_STATIC_SCRATCH_BYTES];

namespace qs {
namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};
}  // namespace

int mps_scratch_setup(int device) {
    QS_REQUIRE(device >= 0 && device < kScratchDevices, QS_E_NODEV,
               "device %d beyond the %d devices the scratch table holds", device,
               kScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return QS_OK;
    int cur = -1;
    QS_HIP(hipGetDevice(&cur));
    QS_REQUIRE(cur == device, QS_E_ARG, "scratch set-up for device %d while device %d is current",
               device, cur);
    void* ptr = nullptr;
    QS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    QS_REQUIRE(ptr != nullptr, QS_E_HIP, "no device address for the scratch array");
    g_scratch_ptr[device] = ptr;
    return QS_OK;
}

namespace {
// Device address of the current device's scratch array (set up on first use).
int scratch_ptr(char** out) {
    int dev = -1;
    QS_HIP(hipGetDevice(&dev));
    QS_REQUIRE(dev >= 0 && dev < kScratchDevices, QS_E_NODEV, "device %d out of range", dev);
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        p = g_scratch_ptr[dev];
    }
    if (!p) {
        const int st = mps_scratch_setup(dev);
        if (st) return st;
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        p = g_scratch_ptr[dev];
    }
    *out = static_cast<char*>(p);
    return QS_OK;
}
}  // namespace
}  // namespace qs

extern "C" {

int qs_mps_max_chi(void) { return kMaxChi; }

size_t qs_mps_workspace_bytes(int ntasks, int chi_max) {
    if (ntasks <= 0 || chi_max <= 0) return 0;
    const size_t head = sv_blob_offset((size_t)ntasks * sizeof(DevTask));
    return head + (size_t)ntasks * theta_slot_elems(chi_max) * sizeof(cd2);
}

namespace {
int chi_class(int chi) {
    return chi <= 8 ? 8 : chi <= 16 ? 16 : chi <= 32 ? 32 : chi <= 64 ? 64 : 128;
}

// Team widths instantiated per register size class (index 0..3 for chi 8, 16, 32, 64):
// at most 1024 threads (TL * CHI) and at most 16 rows per lane (2 CHI / TL).
constexpr int kLaneChoices[4][4] = {{4, 8, 16, 0}, {4, 8, 16, 32}, {4, 8, 16, 32}, {8, 16, 0, 0}};
std::atomic<int> g_svd_lanes[4] = {{16}, {16}, {16}, {16}};

int class_index(int cls) { return cls == 8 ? 0 : cls == 16 ? 1 : cls == 32 ? 2 : cls == 64 ? 3 : -1; }

bool lanes_allowed(int ci, int tl) {
    for (int c : kLaneChoices[ci])
        if (c != 0 && c == tl) return true;
    return false;
}

#define QS_SVD(C, L)                                                                  \
    hipLaunchKernelGGL((mps_svd_kernel<C, L>), dim3(n), dim3((L) * (C)), 0, hs, t, theta, \
                       slot, status)

void launch_svd(int cls, const DevTask* t, int n, const cd2* theta, size_t slot,
                int32_t* status, hipStream_t hs) {
    const int ci = class_index(cls);
    const int tl = ci >= 0 ? g_svd_lanes[ci].load(std::memory_order_relaxed) : 0;
    switch (cls) {
        case 8:
            if (tl == 4) QS_SVD(8, 4); else if (tl == 8) QS_SVD(8, 8); else QS_SVD(8, 16);
            break;
        case 16:
            if (tl == 4) QS_SVD(16, 4); else if (tl == 8) QS_SVD(16, 8);
            else if (tl == 32) QS_SVD(16, 32); else QS_SVD(16, 16);
            break;
        case 32:
            if (tl == 4) QS_SVD(32, 4); else if (tl == 8) QS_SVD(32, 8);
            else if (tl == 32) QS_SVD(32, 32); else QS_SVD(32, 16);
            break;
        case 64:
            if (tl == 8) QS_SVD(64, 8); else QS_SVD(64, 16);
            break;
        default:
            hipLaunchKernelGGL(mps_svd_big_kernel, dim3(n), dim3(kBigThreads), 0, hs, t,
                               const_cast<cd2*>(theta), slot, status);
            break;
    }
}
#undef QS_SVD
}  // namespace

int qs_mps_svd_lane_choices(int chi_class, int* out, int cap, int* count) {
    clear_error();
    const int ci = class_index(chi_class);
    QS_REQUIRE(ci >= 0, QS_E_ARG, "chi class %d has no team-width choices (8, 16, 32 or 64)",
               chi_class);
    QS_REQUIRE(count != nullptr && cap >= 0 && (cap == 0 || out != nullptr), QS_E_ARG,
               "bad output arguments");
    int k = 0;
    for (int c : kLaneChoices[ci]) {
        if (c == 0) continue;
        if (k < cap) out[k] = c;
        ++k;
    }
    *count = k;
    return QS_OK;
}

int qs_mps_svd_lanes(int chi_class) {
    clear_error();
    const int ci = class_index(chi_class);
    QS_REQUIRE(ci >= 0, QS_E_ARG, "chi class %d has no team width", chi_class);
    return g_svd_lanes[ci].load(std::memory_order_relaxed);
}

int qs_mps_set_svd_lanes(int chi_class, int lanes) {
    clear_error();
    const int ci = class_index(chi_class);
    QS_REQUIRE(ci >= 0, QS_E_ARG, "chi class %d has no team width", chi_class);
    QS_REQUIRE(lanes_allowed(ci, lanes), QS_E_ARG, "team width %d not built for chi class %d",
               lanes, chi_class);
    g_svd_lanes[ci].store(lanes, std::memory_order_relaxed);
    return QS_OK;
}

namespace {
constexpr int kClasses[5] = {8, 16, 32, 64, 128};

// Kernel timing (qs_mps_time_kernels): events from a per-thread pool bracket every
// contraction and SVD launch of a batch on its stream, with the launch's kind and tasks.
struct MpsTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<int32_t> kind, tasks;
    size_t used = 0;
    ~MpsTimer() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
};
thread_local MpsTimer g_mps_timer;

// Records the start event of a timed launch (or nothing when timing is off).
int timer_begin(hipStream_t hs, hipEvent_t* e1) {
    *e1 = nullptr;
    if (!g_mps_timer.on) return QS_OK;
    hipEvent_t e0 = g_mps_timer.next();
    *e1 = g_mps_timer.next();
    QS_REQUIRE(e0 && *e1, QS_E_HIP, "hipEventCreate failed");
    QS_HIP(hipEventRecord(e0, hs));
    return QS_OK;
}

int timer_end(hipStream_t hs, hipEvent_t e1, int kind, int ntasks) {
    if (!e1) return QS_OK;
    QS_HIP(hipEventRecord(e1, hs));
    g_mps_timer.kind.push_back(kind);
    g_mps_timer.tasks.push_back(ntasks);
    return QS_OK;
}

// Validates `tasks` and writes them into `out` grouped by SVD size class (stable within
// a class), with each task's original index; counts[c] = tasks of class kClasses[c].
int plan_batch(const qs_mps_task* tasks, int ntasks, DevTask* out, int32_t* counts,
               int* chi_out) {
    QS_REQUIRE(tasks != nullptr, QS_E_ARG, "null tasks");
    int chi = 1;
    for (int i = 0; i < ntasks; ++i) {
        const qs_mps_task& t = tasks[i];
        QS_REQUIRE(t.sites && t.bond && t.dlog, QS_E_ARG, "task %d: null pointer", i);
        QS_REQUIRE(t.chi_max >= 1, QS_E_ARG, "task %d: chi_max %d < 1", i, t.chi_max);
        QS_REQUIRE(t.chi_max <= kMaxChi, QS_E_CHI, "task %d: chi_max %d > %d", i, t.chi_max,
                   kMaxChi);
        QS_REQUIRE(t.site_stride >= (int64_t)2 * t.chi_max * t.chi_max, QS_E_ARG,
                   "task %d: site_stride %lld < 2*chi_max^2", i, (long long)t.site_stride);
        QS_REQUIRE(t.n_sites >= 2, QS_E_ARG, "task %d: n_sites %d < 2", i, t.n_sites);
        QS_REQUIRE(t.site >= 0 && t.site + 1 < t.n_sites, QS_E_ARG,
                   "task %d: bond (%d, %d) outside the chain of %d sites", i, t.site,
                   t.site + 1, t.n_sites);
        QS_REQUIRE(std::isfinite(t.cutoff) && t.cutoff >= 0, QS_E_ARG, "task %d: bad cutoff", i);
        if (t.chi_max > chi) chi = t.chi_max;
    }
    {
        // No two tasks of one chain may touch the same site.
        std::vector<std::pair<std::pair<const void*, int>, int>> keys(ntasks);
        for (int i = 0; i < ntasks; ++i) keys[i] = {{tasks[i].sites, tasks[i].site}, i};
        std::sort(keys.begin(), keys.end());
        for (int i = 1; i < ntasks; ++i)
            if (keys[i].first.first == keys[i - 1].first.first &&
                keys[i].first.second - keys[i - 1].first.second < 2)
                QS_REQUIRE(false, QS_E_ARG, "tasks %d and %d share a site", keys[i - 1].second,
                           keys[i].second);
    }
    int k = 0;
    for (int c = 0; c < 5; ++c) {
        counts[c] = 0;
        for (int i = 0; i < ntasks; ++i)
            if (chi_class(tasks[i].chi_max) == kClasses[c]) {
                std::memset(&out[k], 0, sizeof(DevTask));
                out[k].t = tasks[i];
                out[k].orig = i;
                ++counts[c];
                ++k;
            }
    }
    *chi_out = chi;
    return QS_OK;
}

// Launches a planned batch whose table is already in device memory: the contraction
// (MFMA up to chi 64 when fp32 is asked, FP64 otherwise and for the chi-128 class,
// which is grouped last), then one SVD launch per size class.  No host transfers.
int launch_batch(const DevTask* dtasks, int ntasks, const int32_t* counts, int chi,
                 int precision, char* ws, int32_t* status_dev, hipStream_t hs) {
    cd2* theta = reinterpret_cast<cd2*>(ws + sv_blob_offset((size_t)ntasks * sizeof(DevTask)));
    const size_t slot = theta_slot_elems(chi);
    const int nreg = ntasks - counts[4];
    hipEvent_t e1 = nullptr;
    int st = QS_OK;
    if (precision == QS_PREC_FP32_MFMA && nreg > 0) {
        if ((st = timer_begin(hs, &e1))) return st;
        hipLaunchKernelGGL(mps_contract_f32_mfma_kernel, dim3(nreg), dim3(kMfmaThreads), 0, hs,
                           dtasks, theta, slot, status_dev);
        if ((st = timer_end(hs, e1, 0, nreg))) return st;
    }
    if (precision == QS_PREC_FP32_MFMA && counts[4] > 0) {
        if ((st = timer_begin(hs, &e1))) return st;
        hipLaunchKernelGGL(mps_contract_f32_mfma_big_kernel, dim3(counts[4], 4),
                           dim3(kMfmaThreads), 0, hs, dtasks + nreg,
                           theta + (size_t)nreg * slot, slot, status_dev);
        if ((st = timer_end(hs, e1, 0, counts[4]))) return st;
    }
    const int f64_first = precision == QS_PREC_FP32_MFMA ? ntasks : 0;
    if (ntasks > f64_first) {
        if ((st = timer_begin(hs, &e1))) return st;
        hipLaunchKernelGGL(mps_contract_fp64_kernel, dim3(ntasks - f64_first), dim3(256), 0, hs,
                           dtasks + f64_first, theta + (size_t)f64_first * slot, slot,
                           status_dev);
        if ((st = timer_end(hs, e1, 1, ntasks - f64_first))) return st;
    }
    QS_HIP(hipGetLastError());
    int b = 0;
    for (int c = 0; c < 5; ++c) {
        if (counts[c] == 0) continue;
        if ((st = timer_begin(hs, &e1))) return st;
        launch_svd(kClasses[c], dtasks + b, counts[c], theta + (size_t)b * slot, slot,
                   status_dev, hs);
        if ((st = timer_end(hs, e1, 2, counts[c]))) return st;
        QS_HIP(hipGetLastError());
        b += counts[c];
    }
    return QS_OK;
}

int check_batch_args(int ntasks, int precision, const void* workspace) {
    QS_REQUIRE(ntasks >= 0, QS_E_ARG, "negative task count");
    QS_REQUIRE(workspace != nullptr || ntasks == 0, QS_E_ARG, "null workspace");
    QS_REQUIRE(precision == QS_PREC_FP64 || precision == QS_PREC_FP32_MFMA, QS_E_ARG,
               "unknown precision %d", precision);
    return QS_OK;
}
}  // namespace

int qs_mps_update_batch(const qs_mps_task* tasks, int ntasks, int precision, void* workspace,
                        size_t ws_bytes, int32_t* status_dev, void* stream) {
    clear_error();
    int st = check_batch_args(ntasks, precision, workspace);
    if (st) return st;
    if (ntasks == 0) return QS_OK;
    std::vector<DevTask> dt(ntasks);
    int32_t counts[5];
    int chi = 1;
    st = plan_batch(tasks, ntasks, dt.data(), counts, &chi);
    if (st) return st;
    const size_t need = qs_mps_workspace_bytes(ntasks, chi);
    QS_REQUIRE(ws_bytes >= need, QS_E_RANGE, "workspace %zu bytes < %zu needed", ws_bytes, need);
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    DevTask* dtasks = reinterpret_cast<DevTask*>(ws);
    st = h2d_copy_staged(dtasks, dt.data(), (size_t)ntasks * sizeof(DevTask), hs);
    if (st) return st;
    return launch_batch(dtasks, ntasks, counts, chi, precision, ws, status_dev, hs);
}

int qs_mps_update_batch_static(const qs_mps_task* tasks, int ntasks, int precision,
                               int32_t* status_dev, void* stream) {
    clear_error();
    QS_REQUIRE(ntasks >= 0, QS_E_ARG, "negative task count");
    QS_REQUIRE(precision == QS_PREC_FP64 || precision == QS_PREC_FP32_MFMA, QS_E_ARG,
               "unknown precision %d", precision);
    if (ntasks == 0) return QS_OK;
    // the host table is reused by the calls of one thread (no allocation once it has grown)
    thread_local std::vector<DevTask> dt;
    if (dt.size() < (size_t)ntasks) dt.resize(ntasks);
    int32_t counts[5];
    int chi = 1;
    int st = plan_batch(tasks, ntasks, dt.data(), counts, &chi);
    if (st) return st;
    const size_t need = qs_mps_workspace_bytes(ntasks, chi);
    QS_REQUIRE(need <= (size_t)QS_MPS_STATIC_SCRATCH_BYTES, QS_E_RANGE,
               "batch of %d tasks at chi_max %d needs %zu bytes of scratch; the static scratch "
               "holds %zu (use qs_mps_update_batch with a workspace)", ntasks, chi, need,
               (size_t)QS_MPS_STATIC_SCRATCH_BYTES);
    char* ws = nullptr;
    st = scratch_ptr(&ws);
    if (st) return st;
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    DevTask* dtasks = reinterpret_cast<DevTask*>(ws);
    hipEvent_t copied = nullptr;
    st = h2d_copy(dtasks, dt.data(), (size_t)ntasks * sizeof(DevTask), hs, &copied);
    if (st) return st;
    st = launch_batch(dtasks, ntasks, counts, chi, precision, ws, status_dev, hs);
    const int wt = h2d_copy_wait(copied);
    return st ? st : wt;
}

int qs_mps_time_kernels(int enable) {
    clear_error();
    g_mps_timer.on = enable != 0;
    g_mps_timer.used = 0;
    g_mps_timer.kind.clear();
    g_mps_timer.tasks.clear();
    return QS_OK;
}

int qs_mps_kernel_times(int32_t* kind, int32_t* tasks, double* ms, int cap, int* count) {
    clear_error();
    QS_REQUIRE(count != nullptr && cap >= 0 &&
                   (cap == 0 || (kind != nullptr && tasks != nullptr && ms != nullptr)),
               QS_E_ARG, "bad output arguments");
    const int n = (int)g_mps_timer.kind.size();
    for (int i = 0; i < n && i < cap; ++i) {
        QS_HIP(hipEventSynchronize(g_mps_timer.pool[2 * i + 1]));
        float t = 0.0f;
        QS_HIP(hipEventElapsedTime(&t, g_mps_timer.pool[2 * i], g_mps_timer.pool[2 * i + 1]));
        kind[i] = g_mps_timer.kind[i];
        tasks[i] = g_mps_timer.tasks[i];
        ms[i] = t;
    }
    *count = n;
    g_mps_timer.used = 0;
    g_mps_timer.kind.clear();
    g_mps_timer.tasks.clear();
    return QS_OK;
}

size_t qs_mps_batch_table_bytes(int ntasks) {
    return ntasks > 0 ? (size_t)ntasks * sizeof(DevTask) : 0;
}

int qs_mps_batch_plan(const qs_mps_task* tasks, int ntasks, void* table, size_t table_bytes,
                      int32_t* counts, int* chi_max_out) {
    clear_error();
    QS_REQUIRE(ntasks >= 0, QS_E_ARG, "negative task count");
    QS_REQUIRE(counts != nullptr && chi_max_out != nullptr, QS_E_ARG, "null outputs");
    QS_REQUIRE(ntasks == 0 || table != nullptr, QS_E_ARG, "null table");
    QS_REQUIRE(table_bytes >= qs_mps_batch_table_bytes(ntasks), QS_E_RANGE,
               "table %zu bytes < %zu needed", table_bytes, qs_mps_batch_table_bytes(ntasks));
    if (ntasks == 0) {
        for (int c = 0; c < 5; ++c) counts[c] = 0;
        *chi_max_out = 1;
        return QS_OK;
    }
    return plan_batch(tasks, ntasks, static_cast<DevTask*>(table), counts, chi_max_out);
}

int qs_mps_batch_launch(const void* table_dev, int ntasks, const int32_t* counts, int chi_max,
                        int precision, void* workspace, size_t ws_bytes, int32_t* status_dev,
                        void* stream) {
    clear_error();
    int st = check_batch_args(ntasks, precision, workspace);
    if (st) return st;
    if (ntasks == 0) return QS_OK;
    QS_REQUIRE(table_dev != nullptr && counts != nullptr, QS_E_ARG, "null table or counts");
    QS_REQUIRE(chi_max >= 1 && chi_max <= kMaxChi, QS_E_CHI, "chi_max %d outside [1, %d]",
               chi_max, kMaxChi);
    int sum = 0;
    for (int c = 0; c < 5; ++c) {
        QS_REQUIRE(counts[c] >= 0, QS_E_ARG, "negative class count");
        sum += counts[c];
    }
    QS_REQUIRE(sum == ntasks, QS_E_ARG, "class counts sum to %d, not %d", sum, ntasks);
    for (int c = 0; c < 5; ++c)
        QS_REQUIRE(counts[c] == 0 || kClasses[c] <= chi_class(chi_max), QS_E_ARG,
                   "tasks of chi class %d exceed chi_max %d", kClasses[c], chi_max);
    const size_t need = qs_mps_workspace_bytes(ntasks, chi_max);
    QS_REQUIRE(ws_bytes >= need, QS_E_RANGE, "workspace %zu bytes < %zu needed", ws_bytes, need);
    return launch_batch(static_cast<const DevTask*>(table_dev), ntasks, counts, chi_max,
                        precision, static_cast<char*>(workspace), status_dev,
                        reinterpret_cast<hipStream_t>(stream));
}

int qs_mps_apply_1q_batch(double* sites, const int32_t* bond, int64_t site_stride, int n,
                          const int32_t* site_list, const double* mats, int count,
                          void* stream) {
    clear_error();
    QS_REQUIRE(sites && bond, QS_E_ARG, "null sites or bond array");
    QS_REQUIRE(count >= 0 && (count == 0 || (site_list && mats)), QS_E_ARG,
               "bad gate list (count %d)", count);
    QS_REQUIRE(n >= 1 && site_stride > 0, QS_E_ARG, "bad chain (n %d, stride %lld)", n,
               (long long)site_stride);
    {
        std::vector<int32_t> sorted(site_list, site_list + count);
        std::sort(sorted.begin(), sorted.end());
        for (int i = 0; i < count; ++i) {
            QS_REQUIRE(sorted[i] >= 0 && sorted[i] < n, QS_E_ARG, "site %d outside the chain of %d",
                       sorted[i], n);
            QS_REQUIRE(i == 0 || sorted[i] != sorted[i - 1], QS_E_ARG,
                       "site %d appears twice (compose gates of one site first)", sorted[i]);
        }
    }
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    for (int g0 = 0; g0 < count; g0 += kMax1qPerLaunch) {
        const int k = std::min(kMax1qPerLaunch, count - g0);
        OneSiteBatch b;
        std::memset(&b, 0, sizeof(b));
        for (int j = 0; j < k; ++j) {
            b.g[j].site = site_list[g0 + j];
            std::memcpy(b.g[j].m, mats + 8 * (size_t)(g0 + j), 8 * sizeof(double));
        }
        hipLaunchKernelGGL(mps_apply_1q_batch_kernel, dim3(k), dim3(256), 0, hs,
                           reinterpret_cast<cd2*>(sites), bond, site_stride, b);
        QS_HIP(hipGetLastError());
    }
    return QS_OK;
}

int qs_mps_reset_site(double* sites, const int32_t* bond, int64_t site_stride, int n, int site,
                      void* stream) {
    clear_error();
    QS_REQUIRE(sites && bond, QS_E_ARG, "null sites or bond array");
    QS_REQUIRE(site >= 0 && site < n && site_stride > 0, QS_E_ARG,
               "site %d outside the chain of %d", site, n);
    hipLaunchKernelGGL(mps_reset_site_kernel, dim3(1), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), reinterpret_cast<cd2*>(sites), bond,
                       site_stride, site);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

int qs_mps_apply_1q(double* sites, const int32_t* bond, int64_t site_stride, int site,
                    const double m[8], void* stream) {
    clear_error();
    QS_REQUIRE(sites && bond && m, QS_E_ARG, "null pointer");
    QS_REQUIRE(site >= 0 && site_stride > 0, QS_E_ARG, "bad site/stride");
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(mps_apply_1q_kernel, dim3(16), dim3(256), 0, hs,
                       reinterpret_cast<cd2*>(sites), bond, site_stride, site,
                       mk2(m[0], m[1]), mk2(m[2], m[3]), mk2(m[4], m[5]), mk2(m[6], m[7]));
    QS_HIP(hipGetLastError());
    return QS_OK;
}

}  // extern "C"
