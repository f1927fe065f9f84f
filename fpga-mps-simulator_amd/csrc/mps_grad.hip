// Reverse-mode step of the two-site MPS update on gfx950: given the sites before (A, B) and
// after (A' = U sqrt(S_t), B' = sqrt(S_t) V^H) one update and the cotangents of A', B',
// the cotangents of A and B.  No counterpart in the reference (mps_core.py has no
// differentiation); the forward it differentiates is mps_core.py:277-350 as the update
// kernels run it.  One workgroup per recorded update; the updates of one brick-wall layer
// are independent, so a layer's backward is one launch.
//
// Cotangent convention: dL = Re sum conj(g) dz.  With m = 2 chi_l rows (a, s0), n = 2 chi_r
// columns (s1, d), k kept values:
//   theta = G . (A B)                                  (contraction, gate mixing)
//   D_j = ||A'[:, j]||, U = A' / D, V^H = B' / D,  S_j = Re (U^H theta V)_jj
//   gU = gA' D,  gV = gB'^H D,  gS_j = D_j / (2 S_j) (Re (U^H gA')_jj + Re (gB' V)_jj)
//   Aa = U^H gU,  Bb = V^H gV,  F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eps s_max^4)
//   X = (F o (Aa - Aa^H)) S + S (F o (Bb - Bb^H)) + diag(gS + i Im(Aa_jj) / S_j)
//   W1 = gU - U Aa,  W2 = gV - V Bb
//   g_theta = U (X V^H + S^-1 W2^H) + W1 S^-1 V^H
//   gT = G^H . g_theta,  gA = gT B^H (over q, d),  gB = A^H gT (over a, p)
// The broadened F keeps near-degenerate singular values finite; the renormalisation
// S -> S_t carries no gradient for scale-invariant losses.  All sums run in index order
// (deterministic).
#include <cmath>
#include <vector>

#include "qsim_internal.h"

namespace qs {
namespace {

constexpr int kBwdThreads = 256;
constexpr int kBwdMaxChi = 128;

__device__ __forceinline__ cd2 cmul(cd2 a, cd2 b) {
    return mk2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ cd2 cmadd(cd2 acc, cd2 a, cd2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ cd2 cjmadd(cd2 acc, cd2 a, cd2 b) {   // acc + conj(a) b
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ cd2 conj2(cd2 a) { return mk2(a.x, -a.y); }

// Complex elements one task uses: theta (m x n), Y (k x n), U, thV, gU, W1 (m x k), V, gV,
// W2 (n x k), Aa, Bb, X (k x k), and 3 k reals (S, D, gS) in 2 k complex elements.
__host__ __device__ inline size_t bwd_need(size_t m, size_t n, size_t k) {
    return m * n + k * n + 4 * m * k + 3 * n * k + 3 * k * k + 2 * k;
}
// Work slot per task for tasks of at most chi_max: rows and columns up to 2 chi_max, kept
// values up to chi_max.
__host__ __device__ inline size_t bwd_slot_elems(int chi) {
    const size_t m = 2 * (size_t)chi;
    return bwd_need(m, m, (size_t)chi);
}

struct DevBwdTask {
    qs_mps_bwd_task t;
};

__global__ void __launch_bounds__(kBwdThreads)
mps_update_backward_kernel(const DevBwdTask* __restrict__ tasks, cd2* __restrict__ work,
                           size_t slot, int32_t* __restrict__ status) {
    const qs_mps_bwd_task& T = tasks[blockIdx.x].t;
    const int cl = T.cl, cm = T.cm, cr = T.cr, k = T.k;
    const int m = 2 * cl, n = 2 * cr;
    const int tid = threadIdx.x;
    if (cl < 1 || cm < 1 || cr < 1 || k < 1 || k > m || k > n || cl > kBwdMaxChi ||
        cm > kBwdMaxChi || cr > kBwdMaxChi || bwd_need(m, n, k) > slot) {
        if (tid == 0 && status) status[blockIdx.x] = QS_E_ARG;
        return;
    }
    const cd2* A = reinterpret_cast<const cd2*>(T.a);
    const cd2* B = reinterpret_cast<const cd2*>(T.b);
    const cd2* A2 = reinterpret_cast<const cd2*>(T.a2);
    const cd2* B2 = reinterpret_cast<const cd2*>(T.b2);
    const cd2* gA2 = reinterpret_cast<const cd2*>(T.ga2);
    const cd2* gB2 = reinterpret_cast<const cd2*>(T.gb2);
    cd2* gA = reinterpret_cast<cd2*>(T.ga);
    cd2* gB = reinterpret_cast<cd2*>(T.gb);
    // slot layout (m x n blocks first)
    cd2* th = work + (size_t)blockIdx.x * slot;      // theta, later g_theta, later gT
    cd2* Y = th + (size_t)m * n;                     // k x n
    cd2* U = Y + (size_t)k * n;                      // m x k
    cd2* thV = U + (size_t)m * k;                    // m x k
    cd2* gU = thV + (size_t)m * k;                   // m x k
    cd2* W1 = gU + (size_t)m * k;                    // m x k
    cd2* V = W1 + (size_t)m * k;                     // n x k
    cd2* gV = V + (size_t)n * k;                     // n x k
    cd2* W2 = gV + (size_t)n * k;                    // n x k
    cd2* Aa = W2 + (size_t)n * k;                    // k x k
    cd2* Bb = Aa + (size_t)k * k;                    // k x k
    cd2* X = Bb + (size_t)k * k;                     // k x k
    double* S = reinterpret_cast<double*>(X + (size_t)k * k);
    double* D = S + k;
    double* gS = D + k;
    __shared__ cd2 G[16];
    __shared__ double red[kBwdThreads];
    if (tid < 16) G[tid] = mk2(T.gate[2 * tid], T.gate[2 * tid + 1]);
    __syncthreads();

    // theta[(a, s0), (s1, d)] = sum_pq G[2 s0 + s1][2 p + q] sum_b A[a, p, b] B[b, q, d]
    for (int pr = tid; pr < cl * cr; pr += kBwdThreads) {
        const int a = pr / cr, d = pr - a * cr;
        cd2 t00 = mk2(0, 0), t01 = t00, t10 = t00, t11 = t00;
        const cd2* a0 = A + (size_t)(2 * a) * cm;
        const cd2* a1 = a0 + cm;
        for (int b = 0; b < cm; ++b) {
            const cd2 x0 = a0[b], x1 = a1[b];
            const cd2 y0 = B[(size_t)(2 * b) * cr + d], y1 = B[(size_t)(2 * b + 1) * cr + d];
            t00 = cmadd(t00, x0, y0);
            t01 = cmadd(t01, x0, y1);
            t10 = cmadd(t10, x1, y0);
            t11 = cmadd(t11, x1, y1);
        }
        for (int s = 0; s < 4; ++s) {
            cd2 v = cmul(G[4 * s + 0], t00);
            v = cmadd(v, G[4 * s + 1], t01);
            v = cmadd(v, G[4 * s + 2], t10);
            v = cmadd(v, G[4 * s + 3], t11);
            th[(size_t)(2 * a + (s >> 1)) * n + (s & 1) * cr + d] = v;
        }
    }
    // D_j = ||A'[:, j]||  (A' as m x k)
    for (int j = tid; j < k; j += kBwdThreads) {
        double acc = 0.0;
        for (int i = 0; i < m; ++i) {
            const cd2 x = A2[(size_t)i * k + j];
            acc = fma(x.x, x.x, fma(x.y, x.y, acc));
        }
        D[j] = sqrt(acc);
    }
    __syncthreads();
    // U = A' / D,  V[c, j] = conj(B'[j, c]) / D_j,  gU = gA' D,  gV[c, j] = conj(gB'[j, c]) D_j
    for (int idx = tid; idx < m * k; idx += kBwdThreads) {
        const int j = idx % k;
        const double dj = D[j], inv = dj > 0.0 ? 1.0 / dj : 0.0;
        const cd2 x = A2[idx], g = gA2[idx];
        U[idx] = mk2(x.x * inv, x.y * inv);
        gU[idx] = mk2(g.x * dj, g.y * dj);
    }
    for (int idx = tid; idx < n * k; idx += kBwdThreads) {
        const int c = idx / k, j = idx - c * k;
        const double dj = D[j], inv = dj > 0.0 ? 1.0 / dj : 0.0;
        const cd2 x = B2[(size_t)j * n + c], g = gB2[(size_t)j * n + c];
        V[idx] = mk2(x.x * inv, -x.y * inv);
        gV[idx] = mk2(g.x * dj, -g.y * dj);
    }
    __syncthreads();
    // thV = theta V (m x k)
    for (int idx = tid; idx < m * k; idx += kBwdThreads) {
        const int i = idx / k, j = idx - i * k;
        cd2 acc = mk2(0, 0);
        for (int c = 0; c < n; ++c) acc = cmadd(acc, th[(size_t)i * n + c], V[(size_t)c * k + j]);
        thV[idx] = acc;
    }
    __syncthreads();
    // S_j = Re (U^H theta V)_jj;  gS_j = D_j / (2 S_j) (Re (U^H gA')_jj + Re (gB' V)_jj)
    for (int j = tid; j < k; j += kBwdThreads) {
        double s = 0.0, ga = 0.0, gb = 0.0;
        for (int i = 0; i < m; ++i) {
            const cd2 u = U[(size_t)i * k + j];
            const cd2 t = thV[(size_t)i * k + j];
            const cd2 g = gA2[(size_t)i * k + j];
            s = fma(u.x, t.x, fma(u.y, t.y, s));
            ga = fma(u.x, g.x, fma(u.y, g.y, ga));
        }
        for (int c = 0; c < n; ++c) {
            const cd2 g = gB2[(size_t)j * n + c], v = V[(size_t)c * k + j];
            gb = fma(g.x, v.x, fma(-g.y, v.y, gb));
        }
        S[j] = s;
        gS[j] = s != 0.0 ? D[j] / (2.0 * s) * (ga + gb) : 0.0;
    }
    __syncthreads();
    // Aa = U^H gU, Bb = V^H gV (k x k)
    for (int idx = tid; idx < k * k; idx += kBwdThreads) {
        const int i = idx / k, j = idx - i * k;
        cd2 a = mk2(0, 0), b = mk2(0, 0);
        for (int r = 0; r < m; ++r) a = cjmadd(a, U[(size_t)r * k + i], gU[(size_t)r * k + j]);
        for (int c = 0; c < n; ++c) b = cjmadd(b, V[(size_t)c * k + i], gV[(size_t)c * k + j]);
        Aa[idx] = a;
        Bb[idx] = b;
    }
    // s_max^4 (fixed order)
    double part = 0.0;
    for (int j = tid; j < k; j += kBwdThreads) part = fmax(part, S[j] * S[j]);
    red[tid] = part;
    __syncthreads();
    for (int o = kBwdThreads / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
        __syncthreads();
    }
    const double smax4 = red[0] * red[0];
    // X, W1 = gU - U Aa, W2 = gV - V Bb
    for (int idx = tid; idx < k * k; idx += kBwdThreads) {
        const int i = idx / k, j = idx - i * k;
        cd2 x = mk2(0, 0);
        if (i != j) {
            const double diff = S[j] * S[j] - S[i] * S[i];
            const double f = diff / fma(diff, diff, T.eps * smax4);
            const cd2 a = Aa[idx], ah = conj2(Aa[(size_t)j * k + i]);
            const cd2 b = Bb[idx], bh = conj2(Bb[(size_t)j * k + i]);
            x = mk2(f * (a.x - ah.x) * S[j] + S[i] * f * (b.x - bh.x),
                    f * (a.y - ah.y) * S[j] + S[i] * f * (b.y - bh.y));
        } else {
            const double sinv = S[j] > 0.0 ? 1.0 / S[j] : 0.0;
            x = mk2(gS[j], Aa[idx].y * sinv);
        }
        X[idx] = x;
    }
    for (int idx = tid; idx < m * k; idx += kBwdThreads) {
        const int r = idx / k, j = idx - r * k;
        cd2 acc = gU[idx];
        for (int l = 0; l < k; ++l) {
            const cd2 p = cmul(U[(size_t)r * k + l], Aa[(size_t)l * k + j]);
            acc = mk2(acc.x - p.x, acc.y - p.y);
        }
        W1[idx] = acc;
    }
    for (int idx = tid; idx < n * k; idx += kBwdThreads) {
        const int c = idx / k, j = idx - c * k;
        cd2 acc = gV[idx];
        for (int l = 0; l < k; ++l) {
            const cd2 p = cmul(V[(size_t)c * k + l], Bb[(size_t)l * k + j]);
            acc = mk2(acc.x - p.x, acc.y - p.y);
        }
        W2[idx] = acc;
    }
    __syncthreads();
    // Y = X V^H + S^-1 W2^H (k x n); then g_theta = U Y + W1 S^-1 V^H into th.
    for (int idx = tid; idx < k * n; idx += kBwdThreads) {
        const int i = idx / n, c = idx - i * n;
        cd2 acc = mk2(0, 0);
        for (int j = 0; j < k; ++j) acc = cmadd(acc, X[(size_t)i * k + j], conj2(V[(size_t)c * k + j]));
        const double sinv = S[i] > 0.0 ? 1.0 / S[i] : 0.0;
        const cd2 w = conj2(W2[(size_t)c * k + i]);
        Y[idx] = mk2(acc.x + w.x * sinv, acc.y + w.y * sinv);
    }
    __syncthreads();
    for (int idx = tid; idx < m * n; idx += kBwdThreads) {
        const int r = idx / n, c = idx - r * n;
        cd2 acc = mk2(0, 0);
        for (int i = 0; i < k; ++i) {
            acc = cmadd(acc, U[(size_t)r * k + i], Y[(size_t)i * n + c]);
            const double sinv = S[i] > 0.0 ? 1.0 / S[i] : 0.0;
            const cd2 w = W1[(size_t)r * k + i];
            acc = cmadd(acc, mk2(w.x * sinv, w.y * sinv), conj2(V[(size_t)c * k + i]));
        }
        th[idx] = acc;
    }
    __syncthreads();
    // gT[(a, p), (q, d)] = sum_s conj(G[s][2p + q]) g_theta[(a, s0), (s1, d)], in place per (a, d)
    for (int pr = tid; pr < cl * cr; pr += kBwdThreads) {
        const int a = pr / cr, d = pr - a * cr;
        cd2 gt[4];
        for (int s = 0; s < 4; ++s) gt[s] = th[(size_t)(2 * a + (s >> 1)) * n + (s & 1) * cr + d];
        for (int t = 0; t < 4; ++t) {
            cd2 v = mk2(0, 0);
            for (int s = 0; s < 4; ++s) v = cjmadd(v, G[4 * s + t], gt[s]);
            th[(size_t)(2 * a + (t >> 1)) * n + (t & 1) * cr + d] = v;   // reuse: gT
        }
    }
    __syncthreads();
    // gA[a, p, b] = sum_{q, d} gT[(a, p), (q, d)] conj(B[b, q, d])
    for (int idx = tid; idx < cl * 2 * cm; idx += kBwdThreads) {
        const int row = idx / cm, b = idx - row * cm;          // row = 2a + p
        cd2 acc = mk2(0, 0);
        for (int q = 0; q < 2; ++q)
            for (int d = 0; d < cr; ++d)
                acc = cmadd(acc, th[(size_t)row * n + q * cr + d],
                            conj2(B[(size_t)(2 * b + q) * cr + d]));
        gA[idx] = acc;
    }
    // gB[b, q, d] = sum_{a, p} conj(A[a, p, b]) gT[(a, p), (q, d)]
    for (int idx = tid; idx < cm * 2 * cr; idx += kBwdThreads) {
        const int b = idx / (2 * cr), rem = idx - b * 2 * cr;    // rem = q cr + d
        cd2 acc = mk2(0, 0);
        for (int row = 0; row < m; ++row)
            acc = cjmadd(acc, A[(size_t)row * cm + b], th[(size_t)row * n + rem]);
        gB[idx] = acc;
    }
    if (tid == 0 && status) status[blockIdx.x] = QS_OK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

size_t qs_mps_backward_work_bytes(int ntasks, int chi_max) {
    if (ntasks <= 0 || chi_max <= 0) return 0;
    return sv_blob_offset((size_t)ntasks * sizeof(DevBwdTask)) +
           (size_t)ntasks * bwd_slot_elems(chi_max) * sizeof(cd2);
}

int qs_mps_backward_batch(const qs_mps_bwd_task* tasks, int ntasks, int chi_max, void* work,
                          size_t work_bytes, int32_t* status_dev, void* stream) {
    clear_error();
    QS_REQUIRE(ntasks >= 0, QS_E_ARG, "negative task count");
    if (ntasks == 0) return QS_OK;
    QS_REQUIRE(tasks != nullptr && work != nullptr, QS_E_ARG, "null tasks or work buffer");
    QS_REQUIRE(chi_max >= 1 && chi_max <= kBwdMaxChi, QS_E_CHI, "chi_max %d outside [1, %d]",
               chi_max, kBwdMaxChi);
    for (int i = 0; i < ntasks; ++i) {
        const qs_mps_bwd_task& t = tasks[i];
        QS_REQUIRE(t.a && t.b && t.a2 && t.b2 && t.ga2 && t.gb2 && t.ga && t.gb, QS_E_ARG,
                   "task %d: null pointer", i);
        QS_REQUIRE(t.cl >= 1 && t.cm >= 1 && t.cr >= 1 && t.k >= 1, QS_E_ARG,
                   "task %d: dimensions (%d, %d, %d, k %d)", i, t.cl, t.cm, t.cr, t.k);
        QS_REQUIRE(t.cl <= chi_max && t.cm <= chi_max && t.cr <= chi_max && t.k <= chi_max,
                   QS_E_CHI, "task %d: a dimension exceeds chi_max %d", i, chi_max);
        QS_REQUIRE(t.k <= 2 * t.cl && t.k <= 2 * t.cr, QS_E_ARG,
                   "task %d: k %d exceeds the rank of a %d x %d theta", i, t.k, 2 * t.cl,
                   2 * t.cr);
        QS_REQUIRE(std::isfinite(t.eps) && t.eps >= 0, QS_E_ARG, "task %d: bad eps", i);
    }
    const size_t need = qs_mps_backward_work_bytes(ntasks, chi_max);
    QS_REQUIRE(work_bytes >= need, QS_E_RANGE, "work buffer %zu bytes < %zu needed", work_bytes,
               need);
    hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    std::vector<DevBwdTask> dt(ntasks);
    for (int i = 0; i < ntasks; ++i) dt[i].t = tasks[i];
    char* ws = static_cast<char*>(work);
    DevBwdTask* dtasks = reinterpret_cast<DevBwdTask*>(ws);
    cd2* slots = reinterpret_cast<cd2*>(ws + sv_blob_offset((size_t)ntasks * sizeof(DevBwdTask)));
    int st = h2d_copy_staged(dtasks, dt.data(), (size_t)ntasks * sizeof(DevBwdTask), hs);
    if (st) return st;
    hipLaunchKernelGGL(mps_update_backward_kernel, dim3(ntasks), dim3(kBwdThreads), 0, hs, dtasks,
                       slots, bwd_slot_elems(chi_max), status_dev);
    QS_HIP(hipGetLastError());
    return QS_OK;
}

}  // extern "C"
